"""FP64 CPU oracle for the ED hot path (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
package, and only as the checker or the CPU baseline; the product library
never does. ``liboracle.so`` is the C restatement in qmb_oracle.c (built by
this directory's Makefile); ``model`` restates the host-side model building.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i16p = np.ctypeslib.ndpointer(dtype=np.int16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    c = ctypes
    L.orc_binom.restype = c.c_uint64
    L.orc_binom.argtypes = [c.c_int, c.c_int]
    L.orc_rank_ref.restype = c.c_int64
    L.orc_rank_ref.argtypes = [c.c_uint64, c.c_int]
    L.orc_rank_colex.restype = c.c_int64
    L.orc_rank_colex.argtypes = [c.c_uint64]
    L.orc_unrank_colex.restype = c.c_uint64
    L.orc_unrank_colex.argtypes = [c.c_int64, c.c_int]
    L.orc_sector_states.restype = c.c_int64
    L.orc_sector_states.argtypes = [c.c_int, c.c_int, _u64p]
    L.orc_apply.restype = c.c_uint64
    L.orc_apply.argtypes = [c.c_uint64, _u32p, c.c_int, c.c_int]
    L.orc_rank_colex_many.restype = None
    L.orc_rank_colex_many.argtypes = [_u64p, c.c_int64, _i64p]
    L.orc_srs_min.restype = None
    L.orc_srs_min.argtypes = [c.c_int, c.c_int, _u32p, c.c_void_p, _u64p, c.c_int64, _i64p]
    L.orc_sort_unique.restype = c.c_int64
    L.orc_sort_unique.argtypes = [_i64p, c.c_int64]
    L.orc_orbit_tables.restype = c.c_int
    L.orc_orbit_tables.argtypes = [c.c_int, c.c_int, _u32p, c.c_void_p, c.c_int, _f64p, _u64p, c.c_int64,
                                   _i64p, c.c_int64, _i32p, _f64p, _i64p]
    L.orc_build_orbits.restype = c.c_int64
    L.orc_build_orbits.argtypes = [c.c_int, c.c_int, c.c_int, _u32p, c.c_void_p, c.c_int, _f64p,
                                   _u64p, c.c_int64, _i64p, _i32p, _f64p, _i64p]
    L.orc_closed_form_norms.restype = None
    L.orc_closed_form_norms.argtypes = [c.c_int64, c.c_int, c.c_int, _i32p, _f64p, c.c_double, _f64p]
    L.orc_ell_build.restype = c.c_int
    L.orc_ell_build.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, _f64p, c.c_int, c.c_int, c.c_int,
                                _i32p, _i32p, _f64p, _i32p, _u64p, c.c_int64, _i64p, _f64p, _i64p,
                                c.c_double, c.c_int, c.c_int, _u32p, _f64p, _i16p]
    L.orc_ell_rows_from_reps.restype = c.c_int
    L.orc_ell_rows_from_reps.argtypes = [c.c_int, c.c_int, _u32p, c.c_void_p, c.c_int, _f64p, c.c_int, c.c_int,
                                         c.c_int, _i32p, _i32p, _f64p, _i32p, _u64p, c.c_int64, _i64p, c.c_int64,
                                         c.c_double, c.c_int, c.c_int, _u32p, _f64p, _i16p]
    L.orc_expectation.restype = None
    L.orc_expectation.argtypes = [c.c_int, c.c_int, _u32p, c.c_void_p, c.c_int, _f64p, c.c_int, c.c_int, c.c_int,
                                  _i32p, _i32p, _f64p, _i32p, _u64p, c.c_int64, _i64p, _f64p, _i64p, c.c_double,
                                  c.c_int, _f64p, _f64p]
    L.orc_spmv.restype = None
    L.orc_spmv.argtypes = [c.c_int64, c.c_int, _u32p, _f64p, _i16p, c.c_int, _f64p, _f64p]
    L.orc_start_vector.restype = None
    L.orc_start_vector.argtypes = [c.c_uint64, c.c_int64, c.c_int, _f64p]
    L.orc_lanczos.restype = c.c_int
    L.orc_lanczos.argtypes = [c.c_int64, c.c_int, _u32p, _f64p, _i16p, c.c_int, c.c_int, c.c_uint64,
                              c.c_void_p, _f64p, _f64p]
    L.orc_lanczos_converge.restype = c.c_int
    L.orc_lanczos_converge.argtypes = [c.c_int64, c.c_int, _u32p, _f64p, _i16p, c.c_int, c.c_int, c.c_int,
                                       c.c_double, c.c_uint64, c.c_void_p, _f64p, _f64p,
                                       c.POINTER(c.c_int), c.POINTER(c.c_double)]
    L.orc_lanczos_eigvec.restype = c.c_double
    L.orc_lanczos_eigvec.argtypes = [c.c_int64, c.c_int, _u32p, _f64p, _i16p, c.c_int, c.c_int, c.c_uint64,
                                     c.c_void_p, _f64p, c.c_double, _f64p]
    L.orc_tridiag_ql.restype = c.c_int
    L.orc_tridiag_ql.argtypes = [c.c_int, _f64p, _f64p, _f64p, c.c_void_p]
    L.orc_num_threads.restype = c.c_int
    L.orc_num_threads.argtypes = []
    _LIB = L
    return L


def binom(n, k):
    return int(lib().orc_binom(n, k))


def rank_ref(state, n_up):
    return int(lib().orc_rank_ref(int(state), n_up))


def rank_colex(state):
    return int(lib().orc_rank_colex(int(state)))


def unrank_colex(rank, n_up):
    return int(lib().orc_unrank_colex(int(rank), n_up))


def sector_states(n_sites, n_up):
    out = np.zeros(binom(n_sites, n_up), dtype=np.uint64)
    n = lib().orc_sector_states(n_sites, n_up, out)
    assert n == len(out)
    return out


def apply_elem(state, perm_g, flip, n_sites):
    return int(lib().orc_apply(int(state), np.ascontiguousarray(perm_g, dtype=np.uint32), int(flip), n_sites))


def _chars_f64(chars):
    ch = np.ascontiguousarray(chars, dtype=np.complex128)
    return np.ascontiguousarray(ch.view(np.float64).reshape(-1))


class Orbits:
    """Reference-form orbit data: representatives (by sector rank), stabiliser
    sizes, double-sum norms, closed-form norms and the rank -> (orbit, g) map."""

    def __init__(self, n_sites, n_up, group_el, chars, flip=None, tol_norm=1e-5):
        G, LS = group_el.shape
        assert LS == n_sites
        self.n_sites, self.n_up, self.G = n_sites, n_up, G
        self.n_irreps = chars.shape[0]
        self.group_el = np.ascontiguousarray(group_el, dtype=np.uint32).reshape(-1)
        self.flip = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
        self.chars = _chars_f64(chars)
        self.bas = sector_states(n_sites, n_up)
        C = len(self.bas)
        rep = np.zeros(C, np.int64)
        stab = np.zeros(C, np.int32)
        ns = np.zeros(C * self.n_irreps, np.float64)
        oidx = np.zeros(2 * C, np.int64)
        fl = None if self.flip is None else self.flip.ctypes.data
        n_orb = lib().orc_build_orbits(n_sites, n_up, G, self.group_el, fl, self.n_irreps, self.chars,
                                       self.bas, C, rep, stab, ns, oidx)
        if n_orb < 0:
            raise ValueError("oracle orbit build rejected the input")
        self.n_orb = int(n_orb)
        self.rep_rank = rep[: self.n_orb].copy()
        self.stab = stab[: self.n_orb].copy()
        self.norm_sum = ns[: self.n_orb * self.n_irreps].copy()
        self.orbit_idx = oidx
        self.tol_norm = tol_norm
        self.norms = np.zeros(self.n_orb * self.n_irreps, np.float64)
        lib().orc_closed_form_norms(self.n_orb, G, self.n_irreps, self.stab, self.norm_sum, tol_norm, self.norms)

    @classmethod
    def from_representatives(cls, n_sites, n_up, group_el, chars, flip, bas, rep_rank, tol_norm=1e-5):
        """Orbit data for given sorted representative ranks: the reference's
        SetUpSRS_from_indices + ComputeSRSNorms + ComputeOrbitIndices only
        (used by the CPU baseline to time those steps on the full sector)."""
        self = cls.__new__(cls)
        G, LS = group_el.shape
        self.n_sites, self.n_up, self.G = n_sites, n_up, G
        self.n_irreps = chars.shape[0]
        self.group_el = np.ascontiguousarray(group_el, dtype=np.uint32).reshape(-1)
        self.flip = None if flip is None else np.ascontiguousarray(flip, dtype=np.uint8)
        self.chars = _chars_f64(chars)
        self.bas = bas
        self.rep_rank = np.ascontiguousarray(rep_rank, np.int64)
        self.n_orb = len(self.rep_rank)
        self.stab = np.zeros(self.n_orb, np.int32)
        self.norm_sum = np.zeros(self.n_orb * self.n_irreps, np.float64)
        self.orbit_idx = np.zeros(2 * len(bas), np.int64)
        fl = None if self.flip is None else self.flip.ctypes.data
        rc = lib().orc_orbit_tables(n_sites, G, self.group_el, fl, self.n_irreps, self.chars, bas, len(bas),
                                    self.rep_rank, self.n_orb, self.stab, self.norm_sum, self.orbit_idx)
        if rc:
            raise MemoryError("oracle orbit tables")
        self.tol_norm = tol_norm
        self.norms = np.zeros(self.n_orb * self.n_irreps, np.float64)
        lib().orc_closed_form_norms(self.n_orb, G, self.n_irreps, self.stab, self.norm_sum, tol_norm, self.norms)
        return self

    @property
    def reps(self):
        return self.bas[self.rep_rank]

    def null_mask(self, alpha):
        return self.norms.reshape(self.n_orb, self.n_irreps)[:, alpha] < self.tol_norm


class Ell:
    """Reference-layout ELL built by the oracle restatement of the device kernel."""

    def __init__(self, orbits: Orbits, term_table, alpha, max_terms, is_complex):
        op, site, scalar, n_ops, h_col = term_table
        o = orbits
        self.n_rows, self.W, self.is_complex = o.n_orb, max_terms, bool(is_complex)
        n = o.n_orb
        self.ind = np.zeros(n * max_terms, np.uint32)
        self.val = np.zeros(n * max_terms * (2 if is_complex else 1), np.float64)
        self.track = np.zeros(n, np.int16)
        sc = np.ascontiguousarray(np.asarray(scalar, np.complex128).view(np.float64))
        rc = lib().orc_ell_build(o.n_sites, o.n_up, o.G, o.n_irreps, o.chars, alpha, len(n_ops), h_col,
                                 np.ascontiguousarray(op, np.int32), np.ascontiguousarray(site, np.int32),
                                 sc, np.ascontiguousarray(n_ops, np.int32), o.bas, n, o.rep_rank, o.norms,
                                 o.orbit_idx, o.tol_norm, max_terms, int(is_complex), self.ind, self.val,
                                 self.track)
        if rc:
            raise ValueError("a row needs more slots than max_terms")

    @property
    def values(self):
        if self.is_complex:
            return self.val.view(np.complex128)
        return self.val

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128 if self.is_complex else np.float64)
        y = np.zeros_like(x)
        lib().orc_spmv(self.n_rows, self.W, self.ind, self.val, self.track, int(self.is_complex),
                       x.view(np.float64), y.view(np.float64))
        return y

    def dense(self):
        from .model import ell_to_dense

        return ell_to_dense(self.n_rows, self.W, self.ind, self.values, self.track)

    def lanczos(self, n_iter, seed, start=None):
        a = np.zeros(n_iter, np.float64)
        b = np.zeros(n_iter + 1, np.float64)
        st = None
        if start is not None:
            start = np.ascontiguousarray(start, dtype=np.complex128 if self.is_complex else np.float64)
            st = start.ctypes.data
        it = lib().orc_lanczos(self.n_rows, self.W, self.ind, self.val, self.track, int(self.is_complex),
                               n_iter, seed, st, a, b)
        return a[:it], b[: it + 1]

    def lanczos_converge(self, max_iter, seed, tol=1e-13, check_every=5, start=None):
        a = np.zeros(max_iter, np.float64)
        b = np.zeros(max_iter + 1, np.float64)
        it = ctypes.c_int(0)
        e0 = ctypes.c_double(0.0)
        st = None
        if start is not None:
            start = np.ascontiguousarray(start, dtype=np.complex128 if self.is_complex else np.float64)
            st = start.ctypes.data
        status = lib().orc_lanczos_converge(self.n_rows, self.W, self.ind, self.val, self.track,
                                            int(self.is_complex), max_iter, check_every, tol, seed, st, a, b,
                                            ctypes.byref(it), ctypes.byref(e0))
        n = it.value
        return status, e0.value, n, a[:n], b[: n + 1]

    def eigvec_residual(self, seed, krylov, e0, start=None):
        x = np.zeros(self.n_rows, np.complex128 if self.is_complex else np.float64)
        kr = np.ascontiguousarray(krylov, np.float64)
        st = None
        if start is not None:
            start = np.ascontiguousarray(start, dtype=x.dtype)
            st = start.ctypes.data
        r = lib().orc_lanczos_eigvec(self.n_rows, self.W, self.ind, self.val, self.track, int(self.is_complex),
                                     len(kr), seed, st, kr, e0, x.view(np.float64))
        return r, x


def locality_order(term_table):
    """The term table with its terms reordered as an ELL built with
    QMB_LOCALITY_SLOTS applies them: terms sorted on (lowest site, then highest
    site) descending, stable, a term without operators first. Pure-Python
    restatement of the rule stated in include/qmb.h; the reference has no such
    order (its slots follow the table, utils.cu:133-161)."""
    op, site, scalar, n_ops, h_col = term_table
    op = np.asarray(op, np.int32).reshape(len(n_ops), h_col)
    site = np.asarray(site, np.int32).reshape(len(n_ops), h_col)
    keys = []
    for h in range(len(n_ops)):
        s = [int(site[h, k]) for k in range(int(n_ops[h]))]
        keys.append((min(s), max(s)) if s else (float("inf"), float("inf")))
    order = sorted(range(len(n_ops)), key=lambda h: (-keys[h][0], -keys[h][1]))
    return (np.ascontiguousarray(op[order].reshape(-1)), np.ascontiguousarray(site[order].reshape(-1)),
            np.ascontiguousarray(np.asarray(scalar)[order]), np.ascontiguousarray(np.asarray(n_ops)[order]), h_col)


def ell_rows_from_reps(n_sites, group_el, flip, chars, term_table, alpha, reps, rows, max_terms, is_complex,
                       tol_norm=1e-5):
    """Rows `rows` of the sector matrix whose rows are the ascending representatives
    `reps` (orc_ell_rows_from_reps): orbit data by brute force over the group
    images, targets by binary search in `reps`. Same layout as Ell / export_rows
    (ind[n*W], val, track). Raises if a listed representative is not an orbit
    minimum or a target's orbit is missing from the list."""
    op, site, scalar, n_ops, h_col = term_table
    G = group_el.shape[0]
    rows = np.ascontiguousarray(rows, np.int64)
    n, W = len(rows), int(max_terms)
    ind = np.zeros(n * W, np.uint32)
    val = np.zeros(n * W * (2 if is_complex else 1), np.float64)
    track = np.zeros(n, np.int16)
    fl = None if flip is None else np.ascontiguousarray(flip, np.uint8)
    sc = np.ascontiguousarray(np.asarray(scalar, np.complex128).view(np.float64))
    reps = np.ascontiguousarray(reps, np.uint64)
    rc = lib().orc_ell_rows_from_reps(int(n_sites), G, np.ascontiguousarray(group_el, np.uint32).reshape(-1),
                                      None if fl is None else fl.ctypes.data, chars.shape[0], _chars_f64(chars),
                                      int(alpha), len(n_ops), h_col, np.ascontiguousarray(op, np.int32),
                                      np.ascontiguousarray(site, np.int32), sc, np.ascontiguousarray(n_ops, np.int32),
                                      reps, len(reps), rows, n, tol_norm, W, int(is_complex), ind, val, track)
    if rc == 1:
        raise ValueError("a row needs more slots than max_terms")
    if rc == 2:
        raise ValueError("a listed representative is not the minimum of its orbit")
    if rc == 3:
        raise ValueError("a target's orbit is missing from the representative list")
    if rc:
        raise ValueError(f"orc_ell_rows_from_reps failed ({rc})")
    return ind, (val.view(np.complex128) if is_complex else val), track


def expectation(orbits: Orbits, term_table, alpha, x):
    """<Psi|A|Psi> of the reference-layout sector vector x (orc_expectation): A
    given as a conjugated term table (op, site, scalar, n_ops, h_col), x real or
    complex over the orbit rows. Returns a complex number."""
    op, site, scalar, n_ops, h_col = term_table
    o = orbits
    cplx = np.iscomplexobj(x)
    xx = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64)
    sc = np.ascontiguousarray(np.asarray(scalar, np.complex128).view(np.float64))
    out = np.zeros(2, np.float64)
    fl = None if o.flip is None else o.flip.ctypes.data
    lib().orc_expectation(o.n_sites, o.G, o.group_el, fl, o.n_irreps, o.chars, alpha, len(n_ops), h_col,
                          np.ascontiguousarray(op, np.int32), np.ascontiguousarray(site, np.int32), sc,
                          np.ascontiguousarray(n_ops, np.int32), o.bas, o.n_orb, o.rep_rank, o.norms,
                          o.orbit_idx, o.tol_norm, int(cplx), xx.view(np.float64), out)
    return complex(out[0], out[1])


def spmv_rows(W, ind, val, track, x, is_complex=False):
    """y = H x over a shard of rows given in the reference layout (global
    column indices into the full-length x)."""
    y = np.zeros(len(track) * (2 if is_complex else 1), np.float64)
    lib().orc_spmv(len(track), W, np.ascontiguousarray(ind, np.uint32),
                   np.ascontiguousarray(np.asarray(val).view(np.float64)), np.ascontiguousarray(track, np.int16),
                   int(is_complex), np.ascontiguousarray(np.asarray(x).view(np.float64)), y)
    return y.view(np.complex128) if is_complex else y


def rank_colex_many(states):
    st = np.ascontiguousarray(states, np.uint64)
    out = np.zeros(len(st), np.int64)
    lib().orc_rank_colex_many(st, len(st), out)
    return out


def srs_min(n_sites, group_el, flip, states):
    """Minimum image rank of every state (SetUpSRS + SolveSRS_min)."""
    G = group_el.shape[0]
    st = np.ascontiguousarray(states, np.uint64)
    out = np.zeros(len(st), np.int64)
    fl = None if flip is None else np.ascontiguousarray(flip, np.uint8)
    lib().orc_srs_min(n_sites, G, np.ascontiguousarray(group_el, np.uint32).reshape(-1),
                      None if fl is None else fl.ctypes.data, st, len(st), out)
    return out


def sort_unique(v):
    v = np.ascontiguousarray(v, np.int64).copy()
    n = lib().orc_sort_unique(v, len(v))
    return v[:n]


def start_vector(seed, n, is_complex=False):
    out = np.zeros(2 * n if is_complex else n, np.float64)
    lib().orc_start_vector(seed, n, int(is_complex), out)
    return out.view(np.complex128) if is_complex else out


def tridiag_eig(d, e, vectors=True):
    d = np.ascontiguousarray(d, np.float64)
    e = np.ascontiguousarray(e, np.float64)
    m = len(d)
    w = np.zeros(m, np.float64)
    z = np.zeros(m * m, np.float64) if vectors else None
    rc = lib().orc_tridiag_ql(m, d, e if len(e) else np.zeros(1), w, None if z is None else z.ctypes.data)
    if rc:
        raise RuntimeError("QL did not converge")
    if vectors:
        return w, z.reshape(m, m).T  # column j = eigenvector j
    return w
