"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Integer results (representatives, stabilisers, null flags, lookups, ELL
columns and track) are compared bit for bit; ELL values and SpMV results too,
because both sides use the same FP64 operation sequence (no FMA contraction).
Eigenvalues must agree within 1e-10 relative with the oracle's converged
Lanczos and with dense diagonalisation where it fits.
"""
import contextlib
import os
import sys

import numpy as np
import pytest

import oracle
import qmbcuda_amd as q
from oracle import model as OM
from tests.helpers import (DeviceBuffer, chain_case, ell_dense_nonnull, oracle_ell, oracle_orbits,
                           square_case)

pytestmark = pytest.mark.gpu

SZ, SP, SM = OM.OP_SZ, OM.OP_SP, OM.OP_SM


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def gpu_objects(ctx, case, n_up, alpha, layout, max_terms=0, value_type=None):
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(case["n_sites"], case["table"])
    b = q.Basis(ctx, g, n_up, alpha, layout)
    e = q.Ell(ctx, b, t, value_type=value_type, max_terms=max_terms)
    return g, t, b, e


CASES = [
    ("chain12_flip", lambda: chain_case(12, 1.0, 1.0, 0.3, flip=True), 6, [0, 5, 12, 17]),
    ("chain16", lambda: chain_case(16), 8, [0, 3, 8]),
    ("chain13_odd", lambda: chain_case(13, 1.0, 0.7, 0.2), 6, [0, 2]),
    ("square4x4_flip", lambda: square_case(4, 4, 1.0, 1.0, flip=True), 8, [0, 5, 16 + 10]),
    ("square4x3", lambda: square_case(4, 3, 1.0, 0.8), 6, [1, 7]),
]


@pytest.mark.parametrize("label,mk,n_up,alphas", CASES)
def test_basis_reference_layout_bit_exact(ctx, label, mk, n_up, alphas):
    case = mk()
    orb = oracle_orbits(case, n_up)
    for a in alphas:
        say(f"basis {label} irrep={a}")
        g = q.Group(case["group_el"], case["chars"], case["flip"])
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        reps, stab, nul = b.export()
        assert b.n_rows == orb.n_orb
        assert np.array_equal(reps, orb.reps)
        assert np.array_equal(stab.astype(np.int32), orb.stab)
        assert np.array_equal(nul.astype(bool), orb.null_mask(a))
        bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
        creps, cstab, cnul = bc.export()
        assert np.array_equal(creps, orb.reps[~orb.null_mask(a)])
        assert not cnul.any()


def test_lookup_matches_orbit_index_map(ctx):
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    orb = oracle_orbits(case, 6)
    say("lookup chain12_flip")
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    b = q.Basis(ctx, g, 6, 0, q.LAYOUT_REFERENCE)
    rows, el = b.lookup(orb.bas)
    assert np.array_equal(rows, orb.orbit_idx[0::2])
    assert np.array_equal(el, orb.orbit_idx[1::2])


@pytest.mark.parametrize("label,mk,n_up,alphas", CASES)
def test_ell_reference_layout_bit_exact(ctx, label, mk, n_up, alphas):
    case = mk()
    orb = oracle_orbits(case, n_up)
    for a in alphas:
        say(f"ell {label} irrep={a}")
        ref = oracle_ell(case, orb, a)
        _, _, b, e = gpu_objects(ctx, case, n_up, a, q.LAYOUT_REFERENCE, max_terms=ref.W,
                                 value_type=q.COMPLEX if ref.is_complex else q.REAL)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)  # bit-exact values
        inf = e.info()
        assert inf.nnz == int(ref.track.sum())
        assert inf.max_row_len == int(ref.track.max())
        # SpMV: same slot order, same rounding -> identical results
        rng = np.random.default_rng(a + 1)
        x = rng.normal(size=orb.n_orb)
        if ref.is_complex:
            x = x + 1j * rng.normal(size=orb.n_orb)
        assert np.array_equal(e.spmv(x), ref.spmv(x))


@contextlib.contextmanager
def value_route(route):
    """'codes': the fill writes value codes (the default); 'collect': the FP64
    fill and the value collection pass (QMB_VALUE_CODES=0)."""
    old = os.environ.get("QMB_VALUE_CODES")
    if route == "collect":
        os.environ["QMB_VALUE_CODES"] = "0"
    else:
        os.environ.pop("QMB_VALUE_CODES", None)
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("QMB_VALUE_CODES", None)
        else:
            os.environ["QMB_VALUE_CODES"] = old


@pytest.mark.parametrize("label,mk,n_up,alphas", CASES)
def test_packed_values_equal_fp64_values(ctx, label, mk, n_up, alphas):
    """A real or complex ELL is stored packed (column | value index words, FP64
    diagonal): its export, SpMV results and eigenvalue are bit-identical to the
    FP64 form, by either packing route (value codes written by the fill, or FP64
    values collected after it), and both routes give the same value tables."""
    case = mk()
    for a in alphas:
        real = bool(np.all(np.asarray(case["chars"])[a].imag == 0))
        for layout in (q.LAYOUT_REFERENCE, q.LAYOUT_COMPACT):
            g = q.Group(case["group_el"], case["chars"], case["flip"])
            t = q.Terms(case["n_sites"], case["table"])
            b = q.Basis(ctx, g, n_up, a, layout)
            ef = q.Ell(ctx, b, t, keep_fp64_values=True)
            jf = ef.info()
            assert jf.packed == 0
            rng = np.random.default_rng(a + 7)
            x = rng.normal(size=b.n_rows)
            if not real:
                x = x + 1j * rng.normal(size=b.n_rows)
            e2, r2, _, _ = q.lanczos(ef, max_iter=300, tol=1e-13)
            tables = []
            for route in ("codes", "collect"):
                say(f"packed {label} irrep={a} layout={layout} route={route}")
                with value_route(route):
                    ep = q.Ell(ctx, b, t)
                ip = ep.info()
                assert ip.packed == 1
                assert 1 <= ip.n_values <= (1024 if real else 2048)
                assert ip.device_bytes < jf.device_bytes
                assert ip.format_bytes < jf.format_bytes == jf.spmv_bytes
                assert ip.spmv_bytes == jf.spmv_bytes and ip.nnz == jf.nnz and ip.stored == jf.stored
                for u, v in zip(ep.export_reference(), ef.export_reference()):
                    assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
                assert np.array_equal(ep.spmv(x), ef.spmv(x))
                e1, r1, _, _ = q.lanczos(ep, max_iter=300, tol=1e-13)
                assert e1[0] == e2[0] and r1.iters == r2.iters
                tables.append((ip.n_values, ip.format_bytes, ip.device_bytes))
            assert tables[0] == tables[1]


def three_body_case(n=12, seed=5):
    """No symmetry (one group element), terms S+_i S-_j Sz_k with a distinct
    random coefficient each: thousands of distinct matrix values, more than a
    packed value table holds."""
    rng = np.random.default_rng(seed)
    ops, sites, scal = [], [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                if len({i, j, k}) < 3:
                    continue
                ops += [SP, SM, SZ]
                sites += [i, j, k]
                scal.append(rng.uniform(0.5, 1.5))
    table = (np.array(ops, np.int32), np.array(sites, np.int32), np.array(scal, np.complex128),
             np.full(len(scal), 3, np.int32), 3)
    return dict(n_sites=n, table=table, group_el=np.arange(n, dtype=np.uint32)[None, :],
                chars=np.ones((1, 1), np.complex128), flip=None)


def test_too_many_values_stay_fp64(ctx):
    """A matrix with more distinct values than a value table holds stays in the
    FP64 form by either packing route (the coded fill is redone with FP64
    values), bit-exact against the oracle's ELL."""
    case = three_body_case()
    n_up = 6
    orb = oracle_orbits(case, n_up)
    W = 400
    ref = oracle_ell(case, orb, 0, max_terms=W, is_complex=False)
    assert len(np.unique(ref.val[ref.val != 0.0])) > 1024
    g = q.Group(case["group_el"], case["chars"], None)
    t = q.Terms(case["n_sites"], case["table"])
    b = q.Basis(ctx, g, n_up, 0, q.LAYOUT_REFERENCE)
    rng = np.random.default_rng(3)
    x = rng.normal(size=orb.n_orb)
    for route in ("codes", "collect"):
        say(f"three-body n=12 route={route}")
        with value_route(route):
            e = q.Ell(ctx, b, t, value_type=q.REAL, max_terms=W)
        assert e.info().packed == 0
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        assert np.array_equal(e.spmv(x), ref.spmv(x))


def chain12_with_constant():
    """The chain12 flip table with an operator-free (constant) term inserted in
    the middle: the locality order puts it first."""
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    op, site, sc, n_ops, mo = case["table"]
    h = len(n_ops) // 2
    op = np.insert(np.asarray(op).reshape(-1, mo), h, np.zeros(mo, np.int32), axis=0).reshape(-1)
    site = np.insert(np.asarray(site).reshape(-1, mo), h, np.zeros(mo, np.int32), axis=0).reshape(-1)
    sc = np.insert(np.asarray(sc), h, 0.3 + 0.0j)
    n_ops = np.insert(np.asarray(n_ops), h, 0)
    return dict(case, table=(op.astype(np.int32), site.astype(np.int32), sc, n_ops.astype(np.int32), mo))


@pytest.mark.parametrize("label,mk,n_up,alphas", [CASES[0], CASES[3], CASES[4],
                                                   ("chain12_constant", chain12_with_constant, 6, [0, 5])])
def test_locality_slots_bit_exact(ctx, label, mk, n_up, alphas):
    """QMB_LOCALITY_SLOTS: the library's reordered slots equal the oracle's ELL
    built from the table in locality order (columns, values, track and SpMV bit
    for bit, FP64 and packed), and the eigenvalue is the table-order one."""
    case = mk()
    loc = dict(case, table=oracle.locality_order(case["table"]))
    orb = oracle_orbits(case, n_up)
    for a in alphas:
        say(f"locality slots {label} irrep={a}")
        ref = oracle_ell(loc, orb, a)
        vt = q.COMPLEX if ref.is_complex else q.REAL
        g = q.Group(case["group_el"], case["chars"], case["flip"])
        t = q.Terms(case["n_sites"], case["table"])
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        rng = np.random.default_rng(a + 3)
        x = rng.normal(size=orb.n_orb)
        if ref.is_complex:
            x = x + 1j * rng.normal(size=orb.n_orb)
        for fp64 in (False, True):
            e = q.Ell(ctx, b, t, value_type=vt, max_terms=ref.W, keep_fp64_values=fp64, locality_slots=True)
            ind, val, track = e.export_reference()
            assert np.array_equal(track, ref.track)
            assert np.array_equal(ind, ref.ind)
            assert np.array_equal(val.view(np.float64), ref.val)
            assert np.array_equal(e.spmv(x), ref.spmv(x))
        e_tab = q.Ell(ctx, b, t, value_type=vt, max_terms=ref.W)
        assert not np.array_equal(e_tab.export_reference()[0], ind)
        e1, _, _, _ = q.lanczos(e, max_iter=300, tol=1e-13)
        e2, _, _, _ = q.lanczos(e_tab, max_iter=300, tol=1e-13)
        assert abs(e1[0] - e2[0]) <= 1e-10 * max(1.0, abs(e2[0]))


def test_compact_layout_is_reference_without_null_rows(ctx):
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    orb = oracle_orbits(case, 6)
    for a in (5, 17):
        say(f"compact chain12_flip irrep={a}")
        ref = oracle_ell(case, orb, a)
        E_ref = ell_dense_nonnull(ref, orb, a)
        _, _, b, e = gpu_objects(ctx, case, 6, a, q.LAYOUT_COMPACT, max_terms=ref.W)
        ind, val, track = e.export_reference()
        n = b.n_rows
        E = OM.ell_to_dense(n, ref.W, ind, val, track)
        assert np.array_equal(E, E_ref)


def test_export_rows_is_a_slice_of_the_full_export(ctx):
    case = chain_case(16, 1.0, 0.8, 0.3)
    say("export_rows chain16")
    _, _, b, e = gpu_objects(ctx, case, 8, 3, q.LAYOUT_COMPACT)
    ind, val, track = e.export_reference()
    W = e.info().max_terms
    for row0, k in ((0, 64), (37, 100), (b.n_rows - 5, 5), (b.n_rows, 0)):
        i2, v2, t2 = e.export_rows(row0, k)
        assert np.array_equal(t2, track[row0:row0 + k])
        assert np.array_equal(i2, ind[row0 * W:(row0 + k) * W])
        assert np.array_equal(v2, val[row0 * W:(row0 + k) * W])
    with pytest.raises(q.QMBError):
        e.export_rows(b.n_rows - 3, 4)


def test_max_terms_overflow_is_an_error(ctx):
    case = chain_case(10)
    say("overflow chain10")
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(10, case["table"])
    b = q.Basis(ctx, g, 5, 0, q.LAYOUT_COMPACT)
    with pytest.raises(q.QMBError) as ei:
        q.Ell(ctx, b, t, max_terms=3)
    assert ei.value.status == 1


def test_bad_requests_are_errors(ctx):
    case = chain_case(10, flip=True)
    say("bad requests")
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    with pytest.raises(q.QMBError):
        q.Basis(ctx, g, 4, 0)  # flip needs n_up = N/2
    with pytest.raises(q.QMBError):
        q.Basis(ctx, g, 5, 99)  # irrep out of range
    case = square_case(4, 3)
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(12, case["table"])
    b = q.Basis(ctx, g, 6, 1)
    with pytest.raises(q.QMBError):
        q.Ell(ctx, b, t, value_type=q.REAL)  # complex characters


def test_edge_sectors(ctx):
    say("edge sectors")
    case = chain_case(8)
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(8, case["table"])
    for n_up in (0, 8):
        b = q.Basis(ctx, g, n_up, 0)
        assert b.n_rows == 1
        e = q.Ell(ctx, b, t)
        ev, rep, _, _ = q.lanczos(e, max_iter=5)
        assert abs(ev[0] - 8 * 0.25) < 1e-14  # all aligned: N * Jz/4
    # a sector whose irrep keeps no orbit at all: k=pi for n_up=0 -> zero rows
    b = q.Basis(ctx, g, 0, 4)
    assert b.n_rows == 0


@pytest.mark.parametrize("label,mk,n_up,a", [
    ("chain16_k0", lambda: chain_case(16), 8, 0),
    ("chain14_k3", lambda: chain_case(14, 1.0, 0.8, 0.25), 7, 3),
    ("square4x4_flip_k0", lambda: square_case(4, 4, flip=True), 8, 0),
    ("square4x3_k1", lambda: square_case(4, 3, 1.0, 0.8), 6, 1),
])
def test_lanczos_matches_oracle_and_dense(ctx, label, mk, n_up, a):
    case = mk()
    orb = oracle_orbits(case, n_up)
    say(f"lanczos {label}")
    ref = oracle_ell(case, orb, a)
    st, e_or, it_or, al, be = ref.lanczos_converge(400, 2212456, tol=1e-13, check_every=10)
    assert st == 0
    dense = np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, a))[0]
    _, _, b, e = gpu_objects(ctx, case, n_up, a, q.LAYOUT_COMPACT)
    ev, rep, kr, _ = q.lanczos(e, max_iter=400, tol=1e-13, seed=2212456)
    assert rep.converged
    sc = max(1.0, abs(dense))
    assert abs(ev[0] - e_or) < 1e-10 * sc
    assert abs(ev[0] - dense) < 1e-10 * sc


def test_fixed_iterations_follow_oracle_recurrence(ctx):
    case = chain_case(14, 1.0, 1.0, 0.2)
    orb = oracle_orbits(case, 7)
    say("fixed-iteration recurrence chain14")
    ref = oracle_ell(case, orb, 0)
    a_or, b_or = ref.lanczos(25, 31)
    _, _, b, e = gpu_objects(ctx, case, 7, 0, q.LAYOUT_REFERENCE, max_terms=ref.W)
    ev, rep, kr, _ = q.lanczos(e, max_iter=25, seed=31, fixed_iter=True)
    w, z = oracle.tridiag_eig(a_or, b_or[1:25])
    assert abs(ev[0] - w[0]) < 1e-11
    zr = z[:, 0] * np.sign(z[0, 0])
    np.testing.assert_allclose(kr, zr, atol=1e-8)


def test_known_answers(ctx):
    say("C1 N=16 chain k=0")
    case = chain_case(16)
    _, _, b, e = gpu_objects(ctx, case, 8, 0, q.LAYOUT_COMPACT)
    assert b.n_rows == 810
    ev, rep, _, _ = q.lanczos(e, max_iter=200, tol=1e-13)
    assert abs(ev[0] - (-7.142296360617)) < 1e-11
    say("4x4 square flip-even k=0")
    case = square_case(4, 4, flip=True)
    _, _, b, e = gpu_objects(ctx, case, 8, 0, q.LAYOUT_COMPACT)
    ev, rep, _, _ = q.lanczos(e, max_iter=300, tol=1e-13)
    assert abs(ev[0] - (-11.228483208429)) < 1e-10


def test_c2_full_size_majumdar_ghosh(ctx):
    """N=28 J1-J2 chain at J2/J1 = 1/2, Sz=0, k=0, flip-even: dim 716,728, E0 = -3N/8."""
    import qmbcuda_amd.models as LM

    say("C2 N=28 MG full size")
    lat, info = LM.chain(28, 1.0, 1.0, 0.5)
    ge, fl, ch = LM.translation_group(lat, spin_flip=True)
    g = q.Group(ge, ch, fl)
    t = q.Terms(28, LM.term_table(lat, info))
    b = q.Basis(ctx, g, 14, 0, q.LAYOUT_COMPACT)
    assert b.n_rows == 716728
    e = q.Ell(ctx, b, t)
    inf = e.info()
    assert inf.max_row_len <= 57
    assert inf.packed == 1 and inf.n_values <= 4096
    # size-independent properties: symmetric matrix (<x,Hy> == <Hx,y>)
    rng = np.random.default_rng(0)
    x, y = rng.normal(size=b.n_rows), rng.normal(size=b.n_rows)
    lhs, rhs = x @ e.spmv(y), e.spmv(x) @ y
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)
    ev, rep, _, _ = q.lanczos(e, max_iter=500, tol=1e-13)
    assert rep.converged
    assert abs(ev[0] + 10.5) < 1e-10 * 10.5
    assert rep.n_spmv_timed == 0 and rep.ms_spmv == 0.0
    # sampled SpMV timing (steps 0, 3, 6, ...) does not change the run
    ev3, rep3, _, _ = q.lanczos(e, max_iter=500, tol=1e-13, profile=3)
    assert ev3[0] == ev[0] and rep3.iters == rep.iters
    assert rep3.n_spmv_timed == (rep3.iters + 2) // 3 and rep3.ms_spmv > 0.0
    # the representatives are the orbit minima: spot-check with the lookup
    reps, stab, _ = b.export()
    rows, el = b.lookup(reps[:: 997])
    assert np.array_equal(rows, np.arange(0, b.n_rows, 997))
    assert np.all(el == 0)


def test_lowest_four_with_reorthogonalisation(ctx):
    case = chain_case(14, 1.0, 0.9, 0.3)
    orb = oracle_orbits(case, 7)
    for a in (0, 3):
        say(f"lowest-4 chain14 irrep={a}")
        ref = oracle_ell(case, orb, a)
        dense = np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, a))[:4]
        _, _, b, e = gpu_objects(ctx, case, 7, a, q.LAYOUT_COMPACT)
        ev, rep, _, _ = q.lanczos(e, n_eigs=4, max_iter=300, tol=1e-13)
        assert rep.converged
        np.testing.assert_allclose(ev, dense, atol=1e-10 * max(1, abs(dense[0])))


def test_partial_reorthogonalisation(ctx):
    """reorth = 2 (estimated-loss driven) gives the same lowest eigenvalues as
    dense diagonalisation while orthogonalising on only part of the steps."""
    for label, case, n_up, a in (("chain14", chain_case(14, 1.0, 0.9, 0.3), 7, 3),
                                 ("chain16", chain_case(16, 1.0, 1.0, 0.0), 8, 0)):
        say(f"partial reorthogonalisation {label} irrep={a}")
        orb = oracle_orbits(case, n_up)
        ref = oracle_ell(case, orb, a)
        dense = np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, a))[:3]
        _, _, b, e = gpu_objects(ctx, case, n_up, a, q.LAYOUT_COMPACT)
        ev, rep, _, _ = q.lanczos(e, n_eigs=3, max_iter=400, tol=1e-13, reorth=2)
        print(f"{label}: iterations {rep.iters}, reorthogonalised steps {rep.n_reorth}", flush=True)
        assert rep.converged
        assert rep.n_reorth < rep.iters
        np.testing.assert_allclose(ev, dense, atol=1e-10 * max(1, abs(dense[0])))
        ev1, rep1, _, _ = q.lanczos(e, n_eigs=3, max_iter=400, tol=1e-13, reorth=1)
        assert rep1.n_reorth == rep1.iters
        np.testing.assert_allclose(ev, ev1, atol=1e-10 * max(1, abs(dense[0])))


def test_eigenvector_second_pass(ctx):
    case = chain_case(14, 1.0, 1.0, 0.2)
    orb = oracle_orbits(case, 7)
    say("eigenvector rebuild chain14")
    ref = oracle_ell(case, orb, 0)
    E = ell_dense_nonnull(ref, orb, 0)
    _, _, b, e = gpu_objects(ctx, case, 7, 0, q.LAYOUT_COMPACT)
    ev, rep, kr, x = q.lanczos(e, max_iter=300, tol=1e-13, want_eigvec=True)
    assert abs(np.linalg.norm(x) - 1) < 1e-8
    # the convergence test guarantees a Ritz residual <= sqrt(tol) max(1, |E|); the
    # rebuilt vector's explicit residual must meet it (factor 2 for rounding)
    bound = 2.0 * np.sqrt(1e-13) * max(1.0, abs(ev[0]))
    assert rep.residual[0] <= bound / 2.0
    assert rep.eig_residual <= bound
    assert np.linalg.norm(E @ x - ev[0] * x) <= bound


def test_reference_signature_drop_in(ctx):
    """qmb_lanczos_eigvals_ref on the reference's own row-major arrays (fixed NIter)."""
    case = chain_case(16)
    orb = oracle_orbits(case, 8)
    ref = oracle_ell(case, orb, 0)
    say("eigvals_ref drop-in chain16")
    a_or, b_or = ref.lanczos(50, 2212456)
    w, z = oracle.tridiag_eig(a_or, b_or[1:50])
    E, r, kr = q.eigvals_ref(None, 50, ref.n_rows, ref.W, 2212456, ref.ind, ref.val, ref.track, False, check=True)
    assert abs(E - w[0]) < 1e-11
    res_or, _ = ref.eigvec_residual(2212456, z[:, 0] * np.sign(z[0, 0]), w[0])
    assert abs(r - res_or / ref.n_rows) < 1e-8


def test_reference_signature_device_pointers(ctx):
    """qmb_lanczos_eigvals_ref_device: the reference's free function with its ELL
    arrays as borrowed device pointers gives the host-array drop-in's results bit
    for bit; host pointers and inconsistent arrays are refused."""
    case = chain_case(16)
    orb = oracle_orbits(case, 8)
    ref = oracle_ell(case, orb, 0)
    say("eigvals_ref_device drop-in chain16")
    E_h, r_h, kr_h = q.eigvals_ref(None, 50, ref.n_rows, ref.W, 2212456, ref.ind, ref.val, ref.track, False,
                                   check=True)
    d_ind = DeviceBuffer.from_array(ref.ind)
    d_val = DeviceBuffer.from_array(ref.val)
    d_tr = DeviceBuffer.from_array(ref.track)
    E_d, r_d, kr_d = q.eigvals_ref_device(50, ref.n_rows, ref.W, 2212456, d_ind.ptr, d_val.ptr, d_tr.ptr, False,
                                          check=True)
    assert E_d == E_h and r_d == r_h and np.array_equal(kr_d, kr_h)
    # a host pointer is not device memory
    with pytest.raises(q.QMBError, match="not device memory"):
        q.eigvals_ref_device(50, ref.n_rows, ref.W, 2212456, ref.ind.ctypes.data, d_val.ptr, d_tr.ptr, False)
    # a device array shorter than nobv * max_terms
    short = DeviceBuffer.from_array(ref.ind[: len(ref.ind) // 2])
    if short.reported_size() == (short.ptr, short.nbytes):  # the runtime reports the exact extent
        with pytest.raises(q.QMBError, match="smaller than"):
            q.eigvals_ref_device(50, ref.n_rows, ref.W, 2212456, short.ptr, d_val.ptr, d_tr.ptr, False)
    # a track entry beyond max_terms, then a column beyond nobv
    bad_tr = ref.track.copy()
    bad_tr[3] = ref.W + 1
    d_bad_tr = DeviceBuffer.from_array(bad_tr)
    with pytest.raises(q.QMBError, match="track out of range"):
        q.eigvals_ref_device(50, ref.n_rows, ref.W, 2212456, d_ind.ptr, d_val.ptr, d_bad_tr.ptr, False)
    bad_ind = ref.ind.copy()
    bad_ind[ref.W * 5 + 1] = ref.n_rows + 7
    d_bad_ind = DeviceBuffer.from_array(bad_ind)
    with pytest.raises(q.QMBError, match="column out of range"):
        q.eigvals_ref_device(50, ref.n_rows, ref.W, 2212456, d_bad_ind.ptr, d_val.ptr, d_tr.ptr, False)
    for b_ in (d_ind, d_val, d_tr, short, d_bad_tr, d_bad_ind):
        b_.free()
