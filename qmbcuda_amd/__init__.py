"""Python bindings of libqmb, the MI355X exact-diagonalisation engine.

The product is the C ABI in include/qmb.h (hand-written gfx950 HIP kernels behind
it) and the C++ facade in include/qmb/. This module is a thin ctypes layer used
by the tests and the benchmark; it never computes anything on the CPU in place
of the library, and it raises if libqmb.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import (COMPLEX, EXPECT_GROUP_AVERAGE, KEEP_FP64_VALUES, LAYOUT_COMPACT, LAYOUT_REFERENCE,
                   LOCALITY_SLOTS, REAL, START_RANDOM, START_USER, BasisInfo, EllInfo, LanczosOpts, LanczosReport)
from . import models

__all__ = ["QMBError", "Context", "Group", "Terms", "Basis", "Ell", "lanczos", "lanczos_batched", "models",
           "LAYOUT_REFERENCE", "LAYOUT_COMPACT", "REAL", "COMPLEX", "device_count", "lib"]


class QMBError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        super().__init__(f"{_abi.STATUS_NAMES.get(status, status)}: {message}")


def lib():
    return _abi.load()


def _check(status: int):
    if status != 0:
        raise QMBError(status, lib().qmb_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    return int(lib().qmb_device_count())


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().qmb_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self):
        _check(lib().qmb_ctx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().qmb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Group:
    def __init__(self, group_el, chars, flip=None):
        ge = np.ascontiguousarray(group_el, np.uint32)
        self.G, self.n_sites = ge.shape
        ch = np.ascontiguousarray(chars, np.complex128)
        self.n_irreps = ch.shape[0]
        self.chars = ch
        self._ge = ge
        self._ch = np.ascontiguousarray(ch.view(np.float64))
        self._fl = None if flip is None else np.ascontiguousarray(flip, np.uint8)
        h = C.c_void_p()
        _check(lib().qmb_group_create(self.n_sites, self.G, _ptr(self._ge), _ptr(self._fl), self.n_irreps,
                                      _ptr(self._ch), C.byref(h)))
        self.h = h

    def valid_irreps(self):
        """1 where the character row is a one-dimensional representation of the group."""
        v = np.zeros(self.n_irreps, np.int32)
        _check(lib().qmb_group_irreps_valid(self.h, _ptr(v)))
        return v.astype(bool)

    def info(self):
        inv = np.zeros(self.G, np.int32)
        torus = C.c_int(0)
        _check(lib().qmb_group_info(self.h, _ptr(inv), C.byref(torus)))
        return inv, bool(torus.value)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qmb_group_destroy(self.h)
        except Exception:
            pass


class Terms:
    def __init__(self, n_sites, table):
        op, site, scalar, n_ops, max_ops = table
        self._op = np.ascontiguousarray(op, np.int32)
        self._site = np.ascontiguousarray(site, np.int32)
        self._sc = np.ascontiguousarray(np.asarray(scalar, np.complex128).view(np.float64))
        self._nops = np.ascontiguousarray(n_ops, np.int32)
        self.n_terms = len(self._nops)
        self.scalars = np.asarray(scalar, np.complex128)
        h = C.c_void_p()
        _check(lib().qmb_terms_create(n_sites, self.n_terms, max_ops, _ptr(self._op), _ptr(self._site),
                                      _ptr(self._sc), _ptr(self._nops), C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qmb_terms_destroy(self.h)
        except Exception:
            pass


class Basis:
    def __init__(self, ctx: Context, group: Group, n_up: int, irrep: int, layout=LAYOUT_COMPACT):
        self.ctx, self.group = ctx, group
        h = C.c_void_p()
        _check(lib().qmb_basis_build(ctx.h, group.h, n_up, irrep, layout, C.byref(h)))
        self.h = h
        self.irrep, self.layout = irrep, layout

    def info(self) -> BasisInfo:
        inf = BasisInfo()
        _check(lib().qmb_basis_get_info(self.h, C.byref(inf)))
        return inf

    @property
    def n_rows(self) -> int:
        return int(self.info().n_rows)

    def export(self):
        n = self.n_rows
        reps = np.zeros(n, np.uint64)
        stab = np.zeros(n, np.uint16)
        nul = np.zeros(n, np.uint8)
        _check(lib().qmb_basis_export(self.h, _ptr(reps), _ptr(stab), _ptr(nul)))
        return reps, stab, nul

    def lookup(self, states):
        st = np.ascontiguousarray(states, np.uint64)
        rows = np.zeros(len(st), np.int64)
        el = np.zeros(len(st), np.int32)
        _check(lib().qmb_basis_lookup(self.h, len(st), _ptr(st), _ptr(rows), _ptr(el)))
        return rows, el

    def expectation(self, terms: "Terms", x, group_average: bool = False):
        """<Psi|A|Psi> of the sector vector x (host array, n_rows values, complex
        for a complex irrep); A = terms (a conjugated table). Returns (value, averaged):
        averaged tells whether the group average of A was taken."""
        cplx = np.iscomplexobj(x)
        xx = np.ascontiguousarray(x, np.complex128 if cplx else np.float64)
        if len(xx) != self.n_rows:
            raise ValueError("x must have n_rows entries")
        out = np.zeros(2, np.float64)
        avg = C.c_int(0)
        _check(lib().qmb_expectation(self.ctx.h, self.h, terms.h, int(cplx), _ptr(xx), 0,
                                     EXPECT_GROUP_AVERAGE if group_average else 0, _ptr(out), C.byref(avg)))
        return complex(out[0], out[1]), bool(avg.value)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qmb_basis_destroy(self.h)
        except Exception:
            pass


class Ell:
    def __init__(self, ctx: Context, basis: Basis, terms: Terms, value_type=None, max_terms: int = 0,
                 keep_fp64_values: bool = False, locality_slots: bool = False):
        """keep_fp64_values: store every value in FP64 (no packed value table).
        locality_slots: slots follow the terms sorted on (lowest site, highest
        site) descending instead of table order (QMB_LOCALITY_SLOTS): the same
        matrix, with better gather locality on large sectors."""
        if value_type is None:
            real = models.is_real_irrep(basis.group.chars, basis.irrep) and np.all(terms.scalars.imag == 0)
            value_type = REAL if real else COMPLEX
        self.ctx, self.basis, self.terms = ctx, basis, terms
        self.value_type = value_type
        h = C.c_void_p()
        flags = (KEEP_FP64_VALUES if keep_fp64_values else 0) | (LOCALITY_SLOTS if locality_slots else 0)
        _check(lib().qmb_ell_build(ctx.h, basis.h, terms.h, value_type | flags, max_terms, C.byref(h)))
        self.h = h

    @property
    def is_complex(self) -> bool:
        return self.value_type == COMPLEX

    def save(self, path):
        """Write the device matrix to a file (qmb_ell_save)."""
        _check(lib().qmb_ell_save(self.h, os.fsencode(str(path))))

    @classmethod
    def load(cls, ctx: Context, basis: Basis, path, terms: Terms = None):
        """A matrix saved for this basis (qmb_ell_load)."""
        self = cls.__new__(cls)
        h = C.c_void_p()
        _check(lib().qmb_ell_load(ctx.h, basis.h, os.fsencode(str(path)), C.byref(h)))
        self.h = h
        self.ctx, self.basis, self.terms = ctx, basis, terms
        self.value_type = int(self.info().value_type)
        return self

    def info(self) -> EllInfo:
        inf = EllInfo()
        _check(lib().qmb_ell_get_info(self.h, C.byref(inf)))
        return inf

    def export_reference(self):
        inf = self.info()
        n, W = int(inf.n_rows), int(inf.max_terms)
        ind = np.zeros(n * W, np.uint32)
        val = np.zeros(n * W * (2 if self.is_complex else 1), np.float64)
        track = np.zeros(n, np.int16)
        _check(lib().qmb_ell_export_reference(self.h, _ptr(ind), _ptr(val), _ptr(track)))
        return ind, (val.view(np.complex128) if self.is_complex else val), track

    def export_rows(self, row0, n_rows):
        """Rows [row0, row0 + n_rows) in the reference layout (global column indices)."""
        W = int(self.info().max_terms)
        ind = np.zeros(n_rows * W, np.uint32)
        val = np.zeros(n_rows * W * (2 if self.is_complex else 1), np.float64)
        track = np.zeros(n_rows, np.int16)
        _check(lib().qmb_ell_export_rows(self.h, int(row0), int(n_rows), _ptr(ind), _ptr(val), _ptr(track)))
        return ind, (val.view(np.complex128) if self.is_complex else val), track

    def spmv(self, x):
        x = np.ascontiguousarray(x, np.complex128 if self.is_complex else np.float64)
        y = np.zeros_like(x)
        _check(lib().qmb_spmv_host(self.ctx.h, self.h, _ptr(x), _ptr(y)))
        return y

    def spmv_benchmark(self, reps: int = 20) -> float:
        ms = C.c_double(0.0)
        _check(lib().qmb_spmv_benchmark(self.ctx.h, self.h, reps, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qmb_ell_destroy(self.h)
        except Exception:
            pass


def default_opts(**kw) -> LanczosOpts:
    o = LanczosOpts()
    lib().qmb_lanczos_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def lanczos(ell: Ell, n_eigs=1, max_iter=300, tol=1e-13, seed=2212456, reorth=None, check_every=10,
            fixed_iter=False, start_vec=None, want_eigvec=False, profile=False, allow_unconverged=False):
    """Lowest eigenvalue(s) of an ELL Hamiltonian on the GPU.
    profile: False (off), True (time every SpMV) or k (every k-th) with HIP events;
    report.ms_spmv sums the timed launches, report.n_spmv_timed counts them.
    Returns (evals, report, krylov, eigvec)."""
    if reorth is None:
        reorth = n_eigs > 1
    o = default_opts(max_iter=max_iter, n_eigs=n_eigs, tol=tol, seed=seed, reorth=int(reorth),
                     check_every=check_every, fixed_iter=int(fixed_iter), want_eigvec=int(want_eigvec),
                     profile=int(profile))
    sv = None
    if start_vec is not None:
        sv = np.ascontiguousarray(start_vec, np.complex128 if ell.is_complex else np.float64)
        o.start = START_USER
        o.start_vec = sv.ctypes.data
    ev = np.zeros(n_eigs, np.float64)
    kr = np.zeros(max_iter, np.float64)
    n = ell.basis.n_rows
    vec = np.zeros(n, np.complex128 if ell.is_complex else np.float64) if want_eigvec else None
    rep = LanczosReport()
    st = lib().qmb_lanczos(ell.ctx.h, ell.h, C.byref(o), _ptr(ev), _ptr(kr), _ptr(vec), C.byref(rep))
    if st == 4 and allow_unconverged:
        pass
    else:
        _check(st)
    return ev, rep, kr[: rep.iters], vec


def lanczos_checkpointed(ell: Ell, path, every=0, resume=False, stop_after=0, n_eigs=1, max_iter=300, tol=1e-13,
                         seed=2212456, reorth=None, check_every=10, want_eigvec=False):
    """lanczos with a checkpoint file: written every `every` steps (at a convergence
    check), and when the call stops after `stop_after` steps; resume=True continues
    from it. Returns (evals, report, krylov, eigvec, converged)."""
    if reorth is None:
        reorth = n_eigs > 1
    o = default_opts(max_iter=max_iter, n_eigs=n_eigs, tol=tol, seed=seed, reorth=int(reorth),
                     check_every=check_every, want_eigvec=int(want_eigvec))
    ev = np.zeros(n_eigs, np.float64)
    kr = np.zeros(max_iter, np.float64)
    n = ell.basis.n_rows
    vec = np.zeros(n, np.complex128 if ell.is_complex else np.float64) if want_eigvec else None
    rep = LanczosReport()
    st = lib().qmb_lanczos_checkpointed(ell.ctx.h, ell.h, C.byref(o), os.fsencode(str(path)), int(every),
                                        int(bool(resume)), int(stop_after), _ptr(ev), _ptr(kr), _ptr(vec),
                                        C.byref(rep))
    if st != 4:
        _check(st)
    return ev, rep, kr[: rep.iters], vec, st == 0


def lanczos_batched(ells, n_eigs=1, max_iter=300, tol=1e-13, seed=2212456, reorth=None, check_every=5,
                    profile=False, want_eigvec=False):
    """Solve several ELLs together (one launch per phase for all of them).
    check_every: iterations per graph replay; the host tests convergence after
    each (72 N=24 sectors: 5 is ~3 % faster than 10, 20 ~18 % slower).
    profile: time the graph replays (report.ms_spmv, report.n_spmv_timed).
    want_eigvec: also return every instance's ground-state vector
    (qmb_lanczos_batched_vectors; needs reorth 1 or 2, partial by default).
    Returns (evals[n, n_eigs], reports) or (evals, reports, vectors)."""
    if reorth is None:
        reorth = 2 if want_eigvec else int(n_eigs > 1)
    o = default_opts(max_iter=max_iter, n_eigs=n_eigs, tol=tol, seed=seed, reorth=int(reorth),
                     check_every=check_every, profile=int(profile))
    arr = (C.c_void_p * len(ells))(*[e.h for e in ells])
    ev = np.zeros(len(ells) * n_eigs, np.float64)
    reps = (LanczosReport * len(ells))()
    if not want_eigvec:
        _check(lib().qmb_lanczos_batched(ells[0].ctx.h, len(ells), arr, C.byref(o), _ptr(ev), reps))
        return ev.reshape(len(ells), n_eigs), list(reps)
    vecs = [np.zeros(e.basis.n_rows, np.complex128 if e.is_complex else np.float64) for e in ells]
    vp = (C.c_void_p * len(ells))(*[v.ctypes.data for v in vecs])
    _check(lib().qmb_lanczos_batched_vectors(ells[0].ctx.h, len(ells), arr, C.byref(o), _ptr(ev), vp, reps))
    return ev.reshape(len(ells), n_eigs), list(reps), vecs


class Mlp:
    """Learned start vector network (see include/qmb.h)."""

    def __init__(self, ctx: Context, seed: int = 20261202):
        self.ctx = ctx
        h = C.c_void_p()
        _check(lib().qmb_mlp_create(ctx.h, seed, C.byref(h)))
        self.h = h
        self.n_params = int(lib().qmb_mlp_num_params())
        self._samples = []

    def params(self):
        th = np.zeros(self.n_params, np.float32)
        _check(lib().qmb_mlp_get_params(self.h, _ptr(th)))
        return th

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert len(th) == self.n_params
        _check(lib().qmb_mlp_set_params(self.h, _ptr(th)))

    def forward(self, deltas, j2s):
        d = np.ascontiguousarray(deltas, np.float64)
        j = np.ascontiguousarray(j2s, np.float64)
        out = np.zeros((len(d), 256, 2), np.float32)
        _check(lib().qmb_mlp_forward(self.h, len(d), _ptr(d), _ptr(j), _ptr(out)))
        return out

    def add_sample(self, basis: Basis, ground_state, delta: float, j2: float):
        u = np.ascontiguousarray(ground_state, np.float64)
        _check(lib().qmb_mlp_add_sample(self.h, basis.h, _ptr(u), delta, j2))
        self._samples.append(basis)  # keep the basis alive

    def clear_samples(self):
        _check(lib().qmb_mlp_clear_samples(self.h))
        self._samples = []

    def loss_grad(self, want_grad=True):
        L = C.c_double(0.0)
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        _check(lib().qmb_mlp_loss_grad(self.h, C.byref(L), _ptr(g)))
        return L.value, g

    def train(self, steps: int, lr: float):
        hist = np.zeros(steps, np.float64)
        _check(lib().qmb_mlp_train(self.h, steps, lr, _ptr(hist)))
        return hist

    def set_windows(self, basis: Basis, windows, sublattice_a: int):
        """Register the 8-site windows (array [n_win, 8]) and the Marshall
        sublattice mask of a basis that is not a plain ring."""
        w = np.ascontiguousarray(windows, np.uint8)
        assert w.ndim == 2 and w.shape[1] == 8
        self._keep_windows = getattr(self, "_keep_windows", []) + [basis]
        _check(lib().qmb_mlp_set_windows(self.h, basis.h, int(w.shape[0]), _ptr(w), int(sublattice_a)))

    def start_vector(self, basis: Basis, delta: float, j2: float):
        x = np.zeros(basis.n_rows, np.float64)
        fb = C.c_int(0)
        _check(lib().qmb_mlp_start_vector(self.h, basis.h, delta, j2, _ptr(x), C.byref(fb)))
        return x, bool(fb.value)

    def lanczos(self, basis: Basis, ell: "Ell", delta: float, j2: float, n_eigs=1, max_iter=300, tol=1e-13,
                seed=2212456, check_every=10, want_eigvec=False, profile=False, reorth=None, allow_unconverged=False):
        """Lanczos from the learned start vector kept on the device
        (qmb_lanczos_learned). Returns (evals, report, fallback, eigvec); fallback
        0 = learned run, 1 = unusable learned vector (random start only), 2 = the
        learned run missed the tolerance and the random-start run is the result.
        allow_unconverged: return instead of raising when neither run converged."""
        if reorth is None:
            reorth = n_eigs > 1
        o = default_opts(max_iter=max_iter, n_eigs=n_eigs, tol=tol, seed=seed, reorth=int(reorth),
                         check_every=check_every, want_eigvec=int(want_eigvec), profile=int(profile))
        ev = np.zeros(n_eigs, np.float64)
        vec = np.zeros(basis.n_rows, np.float64) if want_eigvec else None
        rep = LanczosReport()
        fb = C.c_int(0)
        st = lib().qmb_lanczos_learned(self.h, basis.h, ell.h, delta, j2, C.byref(o), _ptr(ev), _ptr(vec),
                                       C.byref(rep), C.byref(fb))
        if not (st == 4 and allow_unconverged):
            _check(st)
        return ev, rep, int(fb.value), vec

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qmb_mlp_destroy(self.h)
        except Exception:
            pass


def eigvals_ref(E_out, NIter, nobv, max_terms, seed, ind, val, track, is_complex, check=True):
    """The reference-signature drop-in (qmb_lanczos_eigvals_ref) on host arrays."""
    E = C.c_double(0.0)
    kr = np.zeros(NIter, np.float64)
    st = C.c_int(0)
    v = np.ascontiguousarray(val.view(np.float64) if is_complex else val, np.float64)
    r = lib().qmb_lanczos_eigvals_ref(C.byref(E), 1, NIter, nobv, max_terms, seed, _ptr(np.ascontiguousarray(ind)),
                                      _ptr(v), _ptr(np.ascontiguousarray(track)), _ptr(kr), 1, int(check),
                                      int(is_complex), C.byref(st))
    _check(st.value)
    return E.value, r, kr


def eigvals_ref_device(NIter, nobv, max_terms, seed, ind_dev, val_dev, track_dev, is_complex, check=True):
    """LanczosEigVals_CUDA with device pointers (qmb_lanczos_eigvals_ref_device):
    ind_dev, val_dev, track_dev are device addresses (ints) of the reference-layout
    arrays (uint32 [nobv*W], float64 [nobv*W] (x2 complex), int16 [nobv])."""
    E = C.c_double(0.0)
    kr = np.zeros(NIter, np.float64)
    st = C.c_int(0)
    r = lib().qmb_lanczos_eigvals_ref_device(C.byref(E), 1, NIter, nobv, max_terms, seed, ind_dev, val_dev, track_dev,
                                             _ptr(kr), 1, int(check), int(is_complex), C.byref(st))
    _check(st.value)
    return E.value, r, kr
