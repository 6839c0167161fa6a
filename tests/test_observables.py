"""Expectation values <Psi|A|Psi> in a symmetry-reduced sector (ComputeStaticExpValZeroT).

The oracle (orc_expectation) is pinned against a dense computation in the full
Sz sector: the sector's rows are embedded as orbit states, the embedding is
checked to carry H's eigenvectors (H Psi = E Psi), and <Psi|A|Psi> is taken with
A built straight from the operator algebra. The GPU kernel is compared with the
oracle on the same vectors (FP64, different summation order: 1e-12), and at full
C2 size through identities that hold exactly.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from oracle import model as OM
from tests.helpers import chain_case, embedding, oracle_orbits, sector_ground_state, square_case

SP, SM, SZ = OM.OP_SP, OM.OP_SM, OM.OP_SZ


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def phase_chain_case(n, theta=0.3, jz=0.8):
    """Chain with a complex hopping phase (Hermitian, translation invariant): its
    eigenvectors are complex in every irrep."""
    geom, _ = OM.chain_xxz_info(n)
    terms = ([[(i, SP, np.exp(1j * theta) / 2), ((i + 1) % n, SM, 1.0 + 0j)] for i in range(n)]
             + [[(i, SM, np.exp(-1j * theta) / 2), ((i + 1) % n, SP, 1.0 + 0j)] for i in range(n)]
             + [[(i, SZ, jz + 0j), ((i + 1) % n, SZ, 1.0 + 0j)] for i in range(n)])
    ge, ch = OM.translation_group(geom)
    return dict(n_sites=n, terms=terms, table=OM.conjugated_term_table(terms), group_el=ge, chars=ch, flip=None)


def op_nonsymmetric():
    """Neither group invariant nor Hermitian."""
    return [[(0, SP, 0.4 + 0.9j), (3, SM, 1.0 + 0j)], [(1, SZ, 1.0 + 0j), (2, SZ, 1.0 + 0j)]]


def op_bonds(n):
    """sum_i S_i . S_{i+1} on a ring (invariant under translations and the spin flip)."""
    t = []
    for i in range(n):
        j = (i + 1) % n
        t += [[(i, SP, 0.5 + 0j), (j, SM, 1.0 + 0j)], [(i, SM, 0.5 + 0j), (j, SP, 1.0 + 0j)],
              [(i, SZ, 1.0 + 0j), (j, SZ, 1.0 + 0j)]]
    return t


CPU_CASES = [
    ("chain8_phase", lambda: phase_chain_case(8), 4),
    ("chain10_flip_j2", lambda: chain_case(10, 1.0, 1.0, 0.4, flip=True), 5),
    ("square4x3", lambda: square_case(4, 3, 1.0, 0.8), 6),
]


def truth(case, orb, alpha, x, terms):
    E = embedding(case, orb, alpha)
    psi = E @ np.conj(x)
    A = OM.dense_sector_hamiltonian(case["n_sites"], orb.n_up, terms)
    return psi.conj() @ A @ psi


@pytest.mark.parametrize("label,mk,n_up", CPU_CASES)
def test_embedding_carries_the_eigenvectors(label, mk, n_up):
    case = mk()
    orb = oracle_orbits(case, n_up)
    H = OM.dense_sector_hamiltonian(case["n_sites"], n_up, case["terms"])
    for a in range(orb.n_irreps):
        if orb.null_mask(a).all():
            continue
        e0, x = sector_ground_state(case, orb, a)
        psi = embedding(case, orb, a) @ np.conj(x)
        assert abs(np.linalg.norm(psi) - 1.0) < 1e-12
        assert np.linalg.norm(H @ psi - e0 * psi) < 1e-10


@pytest.mark.parametrize("label,mk,n_up", CPU_CASES)
def test_oracle_expectation_matches_dense(label, mk, n_up):
    case = mk()
    n = case["n_sites"]
    orb = oracle_orbits(case, n_up)
    ops = [op_nonsymmetric(), op_bonds(n) if label.startswith("chain") else case["terms"]]
    for a in range(orb.n_irreps):
        if orb.null_mask(a).all():
            continue
        e0, x = sector_ground_state(case, orb, a)
        for terms in ops:
            got = oracle.expectation(orb, OM.conjugated_term_table(terms), a, x)
            assert abs(got - truth(case, orb, a, x, terms)) < 1e-12
        # <H> = E0
        assert abs(oracle.expectation(orb, case["table"], a, x) - e0) < 1e-11


def test_oracle_expectation_any_vector():
    """The identity holds for every vector of the sector, not only eigenvectors."""
    case = chain_case(10, 1.0, 1.0, 0.4, flip=True)
    orb = oracle_orbits(case, 5)
    rng = np.random.default_rng(7)
    for a in (0, 3, 10 + 4):
        x = rng.normal(size=orb.n_orb) + 1j * rng.normal(size=orb.n_orb)
        x[orb.null_mask(a)] = 0.0
        for terms in (op_nonsymmetric(), op_bonds(10)):
            got = oracle.expectation(orb, OM.conjugated_term_table(terms), a, x)
            assert abs(got - truth(case, orb, a, x, terms)) < 1e-11


def test_reference_pairing_is_caught():
    """Planted mistake: pairing conj(x_src) ... x_t (the reference kernel's order,
    equivalent to passing conj(x)) gives wrong values for complex vectors, and the
    dense check sees it."""
    case = phase_chain_case(8)
    orb = oracle_orbits(case, 4)
    worst = 0.0
    for a in range(orb.n_irreps):
        e0, x = sector_ground_state(case, orb, a)
        tab = OM.conjugated_term_table(op_nonsymmetric())
        wrong = oracle.expectation(orb, tab, a, np.conj(x))
        worst = max(worst, abs(wrong - truth(case, orb, a, x, op_nonsymmetric())))
    assert worst > 1e-3


def test_operator_builders_match_the_oracle_tables():
    import qmbcuda_amd.models as LM

    t = LM.spin_spin(2, 5, 0.7)
    ref = OM.conjugated_term_table([[(a, b, c) for (a, b, c) in p] for p in t])
    got = LM.operator_table(t)
    for u, v in zip(got[:4], ref[:4]):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    assert got[4] == ref[4]
    # S^zz(k) is A_k^dagger A_k: its dense matrix is the product of the Fourier operators
    lat, _ = LM.chain(6)
    n_up = 3
    Szz = OM.dense_sector_hamiltonian(6, n_up, LM.structure_factor_zz(lat, 2))
    Ak = sum(np.exp(2j * math.pi * 2 * r / 6) / math.sqrt(6) * OM.dense_sector_hamiltonian(6, n_up, [[(r, SZ, 1.0)]])
             for r in range(6))
    np.testing.assert_allclose(Szz, Ak.conj().T @ Ak, atol=1e-13)


# ---------------------------------------------------------------------- GPU

GPU_CASES = [
    ("chain8_phase", lambda: phase_chain_case(8), 4, [0, 1, 3], True),
    ("chain10_flip_j2", lambda: chain_case(10, 1.0, 1.0, 0.4, flip=True), 5, [0, 3, 10 + 5], True),
    ("chain9_odd", lambda: chain_case(9, 1.0, 1.3, 0.2), 4, [0, 2], True),
    ("chain6_small_ring", lambda: chain_case(6, 1.0, 0.9, 0.0), 3, [0, 1], True),
    ("square4x3", lambda: square_case(4, 3, 1.0, 0.8), 6, [0, 1, 7], True),
    ("square4x4_flip", lambda: square_case(4, 4, 1.0, 1.0, flip=True), 8, [0, 5, 16 + 10], False),
]


@pytest.mark.gpu
@pytest.mark.parametrize("label,mk,n_up,alphas,dense_ok", GPU_CASES)
def test_gpu_expectation_matches_oracle(ctx, label, mk, n_up, alphas, dense_ok):
    import qmbcuda_amd as q

    case = mk()
    n = case["n_sites"]
    orb = oracle_orbits(case, n_up)
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    sym = op_bonds(n) if label.startswith("chain") else case["terms"]
    t_ns = q.Terms(n, OM.conjugated_term_table(op_nonsymmetric()))
    t_sym = q.Terms(n, OM.conjugated_term_table(sym))
    rng = np.random.default_rng(11)
    for a in alphas:
        keep = ~orb.null_mask(a)
        if not keep.any():
            continue
        e0, x_gs = sector_ground_state(case, orb, a)
        x_rnd = rng.normal(size=orb.n_orb) + 1j * rng.normal(size=orb.n_orb)
        x_rnd[~keep] = 0.0
        for layout in (q.LAYOUT_REFERENCE, q.LAYOUT_COMPACT):
            b = q.Basis(ctx, g, n_up, a, layout)
            for xname, x in (("gs", x_gs), ("random", x_rnd)):
                xl = x if layout == q.LAYOUT_REFERENCE else x[keep]
                for tname, tg, terms, sym_op in (("nonsym", t_ns, op_nonsymmetric(), False),
                                                 ("sym", t_sym, sym, True)):
                    say(f"expectation {label} irrep={a} layout={layout} x={xname} op={tname}")
                    want = oracle.expectation(orb, OM.conjugated_term_table(terms), a, x)
                    got, averaged = b.expectation(tg, xl)
                    assert averaged == (not sym_op)
                    assert abs(got - want) < 1e-12 * max(1.0, abs(want)), (got, want)
                    if sym_op:
                        got2, averaged2 = b.expectation(tg, xl, group_average=True)
                        assert averaged2
                        assert abs(got2 - want) < 1e-12 * max(1.0, abs(want))
                    if dense_ok and xname == "gs" and layout == q.LAYOUT_COMPACT:
                        assert abs(got - truth(case, orb, a, x, terms)) < 1e-11
        # real vectors on a real irrep take the real path
        if np.all(np.asarray(case["chars"])[a].imag == 0):
            xr = rng.normal(size=int(keep.sum()))
            b = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
            xfull = np.zeros(orb.n_orb)
            xfull[keep] = xr
            say(f"expectation {label} irrep={a} real vector")
            got, _ = b.expectation(t_ns, xr)
            want = oracle.expectation(orb, OM.conjugated_term_table(op_nonsymmetric()), a, xfull)
            assert abs(got - want) < 1e-12 * max(1.0, abs(want))


@pytest.mark.gpu
def test_gpu_expectation_c2_full_size(ctx):
    """N=28 J1-J2 chain (dim 716,728), ground state from the GPU Lanczos:
    <H> = E0; sum_r <Sz_0 Sz_r> = <Sz_0 Sz_total> = 0 in the Sz = 0 sector;
    <Sz_0 Sz_0> = 1/4; S^zz(pi) from the direct operator equals the Fourier sum
    of the correlations (group-averaged path vs invariant path)."""
    import qmbcuda_amd as q
    import qmbcuda_amd.models as LM

    say("expectation C2 N=28 full size")
    lat, info = LM.chain(28, 1.0, 1.0, 0.5)
    ge, fl, ch = LM.translation_group(lat, spin_flip=True)
    g = q.Group(ge, ch, fl)
    th = q.Terms(28, LM.term_table(lat, info))
    b = q.Basis(ctx, g, 14, 0, q.LAYOUT_COMPACT)
    e = q.Ell(ctx, b, th)
    ev, rep, _, vec = q.lanczos(e, max_iter=500, tol=1e-13, want_eigvec=True)
    # the explicit residual of the rebuilt vector meets the convergence test's bound
    assert rep.converged and rep.eig_residual <= 2.0 * np.sqrt(1e-13) * max(1.0, abs(ev[0]))
    nv = float(vec @ vec)
    eh, avg = b.expectation(th, vec)
    assert not avg
    assert abs(eh.real / nv - ev[0]) < 1e-9 and abs(eh.imag) < 1e-12
    corr = []
    for r in range(28):
        v, avg = b.expectation(q.Terms(28, LM.operator_table(LM.szsz(0, r))), vec)
        assert avg
        corr.append(v.real / nv)
    corr = np.array(corr)
    assert abs(corr[0] - 0.25) < 1e-12
    assert abs(corr.sum()) < 1e-11
    s_pi, avg = b.expectation(q.Terms(28, LM.operator_table(LM.structure_factor_zz(lat, 14))), vec)
    assert not avg
    assert abs(s_pi.real / nv - sum(corr[r] * (-1) ** r for r in range(28))) < 1e-10
