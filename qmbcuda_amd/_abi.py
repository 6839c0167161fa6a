"""ctypes declarations of the libqmb C ABI (include/qmb.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqmb.so")

QMB_OK = 0
STATUS_NAMES = {
    0: "QMB_OK",
    1: "QMB_ERR_INVALID",
    2: "QMB_ERR_HIP",
    3: "QMB_ERR_OOM",
    4: "QMB_ERR_NOT_CONVERGED",
    5: "QMB_ERR_NO_DEVICE",
    6: "QMB_ERR_UNSUPPORTED",
}

LAYOUT_REFERENCE, LAYOUT_COMPACT = 0, 1
REAL, COMPLEX = 0, 1
START_RANDOM, START_USER = 0, 1
EXPECT_GROUP_AVERAGE = 1
KEEP_FP64_VALUES = 16
LOCALITY_SLOTS = 32


class BasisInfo(C.Structure):
    _fields_ = [
        ("sector_dim", C.c_int64),
        ("n_rows", C.c_int64),
        ("n_null", C.c_int64),
        ("n_sites", C.c_int32),
        ("n_up", C.c_int32),
        ("irrep", C.c_int32),
        ("layout", C.c_int32),
        ("device_bytes", C.c_int64),
    ]


class EllInfo(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int64),
        ("nnz", C.c_int64),
        ("stored", C.c_int64),
        ("max_row_len", C.c_int32),
        ("max_terms", C.c_int32),
        ("value_type", C.c_int32),
        ("chunk_rows", C.c_int32),
        ("device_bytes", C.c_int64),
        ("spmv_bytes", C.c_double),
        ("packed", C.c_int32),
        ("n_values", C.c_int32),
        ("format_bytes", C.c_double),
    ]


class LanczosOpts(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32),
        ("n_eigs", C.c_int32),
        ("check_every", C.c_int32),
        ("reorth", C.c_int32),
        ("fixed_iter", C.c_int32),
        ("start", C.c_int32),
        ("want_eigvec", C.c_int32),
        ("profile", C.c_int32),
        ("tol", C.c_double),
        ("seed", C.c_uint64),
        ("start_vec", C.c_void_p),
    ]


class LanczosReport(C.Structure):
    _fields_ = [
        ("iters", C.c_int32),
        ("converged", C.c_int32),
        ("breakdown", C.c_int32),
        ("n_spmv", C.c_int32),
        ("n_reorth", C.c_int32),
        ("n_spmv_timed", C.c_int32),
        ("residual", C.c_double * 8),
        ("eig_residual", C.c_double),
        ("ms_total", C.c_double),
        ("ms_spmv", C.c_double),
        ("ms_host", C.c_double),
    ]


# (name, restype, argtypes) of every exported function declared in include/qmb.h
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)
SIGNATURES = [
    ("qmb_last_error", C.c_char_p, []),
    ("qmb_version", C.c_int, []),
    ("qmb_ctx_create", C.c_int, [C.c_int, _pp]),
    ("qmb_ctx_destroy", C.c_int, [_vp]),
    ("qmb_ctx_synchronize", C.c_int, [_vp]),
    ("qmb_device_count", C.c_int, []),
    ("qmb_group_create", C.c_int, [C.c_int, C.c_int, _vp, _vp, C.c_int, _vp, _pp]),
    ("qmb_group_destroy", C.c_int, [_vp]),
    ("qmb_group_info", C.c_int, [_vp, _vp, C.POINTER(C.c_int)]),
    ("qmb_terms_create", C.c_int, [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _pp]),
    ("qmb_terms_destroy", C.c_int, [_vp]),
    ("qmb_basis_build", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _pp]),
    ("qmb_basis_destroy", C.c_int, [_vp]),
    ("qmb_basis_get_info", C.c_int, [_vp, C.POINTER(BasisInfo)]),
    ("qmb_basis_export", C.c_int, [_vp, _vp, _vp, _vp]),
    ("qmb_basis_lookup", C.c_int, [_vp, C.c_int64, _vp, _vp, _vp]),
    ("qmb_ell_build", C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, _pp]),
    ("qmb_ell_destroy", C.c_int, [_vp]),
    ("qmb_ell_get_info", C.c_int, [_vp, C.POINTER(EllInfo)]),
    ("qmb_ell_export_reference", C.c_int, [_vp, _vp, _vp, _vp]),
    ("qmb_ell_export_rows", C.c_int, [_vp, C.c_int64, C.c_int64, _vp, _vp, _vp]),
    ("qmb_group_irreps_valid", C.c_int, [_vp, _vp]),
    ("qmb_ell_save", C.c_int, [_vp, C.c_char_p]),
    ("qmb_ell_load", C.c_int, [_vp, _vp, C.c_char_p, _pp]),
    ("qmb_lanczos_checkpointed", C.c_int, [_vp, _vp, _vp, C.c_char_p, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                           _vp]),
    ("qmb_expectation", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, C.c_int, _vp, C.POINTER(C.c_int)]),
    ("qmb_expectation_noalloc", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, C.POINTER(C.c_int)]),
    ("qmb_spmv_host", C.c_int, [_vp, _vp, _vp, _vp]),
    ("qmb_spmv_device", C.c_int, [_vp, _vp, _vp, _vp]),
    ("qmb_spmv_benchmark", C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_double)]),
    ("qmb_lanczos_default_opts", None, [C.POINTER(LanczosOpts)]),
    ("qmb_lanczos", C.c_int, [_vp, _vp, C.POINTER(LanczosOpts), _vp, _vp, _vp, C.POINTER(LanczosReport)]),
    ("qmb_lanczos_eigvals_ref", C.c_double,
     [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
      C.POINTER(C.c_int)]),
    ("qmb_lanczos_eigvals_ref_device", C.c_double,
     [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int,
      C.POINTER(C.c_int)]),
    ("qmb_lanczos_batched", C.c_int, [_vp, C.c_int, _vp, C.POINTER(LanczosOpts), _vp, _vp]),
    ("qmb_lanczos_batched_vectors", C.c_int, [_vp, C.c_int, _vp, C.POINTER(LanczosOpts), _vp, _vp, _vp]),
    ("qmb_mlp_num_params", C.c_int, []),
    ("qmb_mlp_create", C.c_int, [_vp, C.c_uint64, _pp]),
    ("qmb_mlp_destroy", C.c_int, [_vp]),
    ("qmb_mlp_get_params", C.c_int, [_vp, _vp]),
    ("qmb_mlp_set_params", C.c_int, [_vp, _vp]),
    ("qmb_mlp_forward", C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    ("qmb_mlp_add_sample", C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double]),
    ("qmb_mlp_clear_samples", C.c_int, [_vp]),
    ("qmb_mlp_loss_grad", C.c_int, [_vp, C.POINTER(C.c_double), _vp]),
    ("qmb_mlp_train", C.c_int, [_vp, C.c_int, C.c_double, _vp]),
    ("qmb_mlp_set_windows", C.c_int, [_vp, _vp, C.c_int, _vp, C.c_uint64]),
    ("qmb_mlp_start_vector", C.c_int, [_vp, _vp, C.c_double, C.c_double, _vp, C.POINTER(C.c_int)]),
    ("qmb_lanczos_learned", C.c_int,
     [_vp, _vp, _vp, C.c_double, C.c_double, _vp, _vp, _vp, _vp, C.POINTER(C.c_int)]),
]

_LIB = None


def load():
    """Load libqmb.so (built in-tree by `make -C qmbcuda_amd`). Raises if it is missing:
    there is no CPU fallback."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C qmbcuda_amd` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib
