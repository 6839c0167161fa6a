"""Pins of the FP64 CPU oracle (runs on the CPU, no GPU).

The reference ships no tests, golden vectors or fixtures, so every oracle
function is pinned by an independent check: exact combinatorics, brute-force
orbit enumeration, dense diagonalisation of the full Sz sector built straight
from the operator algebra, closed forms (Majumdar-Ghosh, 2-site, 4-ring) and
known ground-state energies computed with scipy in FP64, plus numpy's
eigh for the tridiagonal solver.
"""
import math
from itertools import combinations

import numpy as np
import pytest

import oracle
from oracle import model as OM
from tests.helpers import (brute_orbits, burnside_count, chain_case, dense_spectrum, ell_dense_nonnull, oracle_ell, oracle_orbits,
                           square_case)


def test_binom_exact():
    for n in range(0, 65):
        for k in range(0, n + 1):
            assert oracle.binom(n, k) == math.comb(n, k)


@pytest.mark.parametrize("n,k", [(10, 5), (12, 4), (9, 3), (14, 7), (16, 8)])
def test_rank_reference_formula_is_colex_position(n, k):
    states = oracle.sector_states(n, k)
    assert len(states) == math.comb(n, k)
    assert np.all(np.diff(states.astype(np.int64)) > 0)
    for pos, s in enumerate(states):
        assert oracle.rank_ref(int(s), k) == pos
        assert oracle.rank_colex(int(s)) == pos
        assert oracle.unrank_colex(pos, k) == int(s)


def test_rank_wide_states():
    # 36-site states (beyond the reference's 32 bits): colex == literal formula on samples
    rng = np.random.default_rng(5)
    for _ in range(200):
        bits = rng.choice(36, 18, replace=False)
        s = sum(1 << int(b) for b in bits)
        assert oracle.rank_ref(s, 18) == oracle.rank_colex(s)
        assert oracle.unrank_colex(oracle.rank_colex(s), 18) == s


def test_sector_states_closed_form():
    for n, k in [(8, 0), (8, 8), (11, 3), (13, 6)]:
        st = oracle.sector_states(n, k)
        expect = sorted(sum(1 << p for p in c) for c in combinations(range(n), k))
        assert st.tolist() == expect


def test_group_convention_chain():
    geom, _ = OM.chain_xxz_info(6)
    ge, ch = OM.translation_group(geom)
    # element 1 = translation by +1; the array maps each site to its preimage
    assert ge[0].tolist() == list(range(6))
    assert ge[1].tolist() == [5, 0, 1, 2, 3, 4]
    # characters: exp(+2 pi i a e / N), exact on quarter turns
    assert ch[0].tolist() == [1] * 6
    assert ch[3, 1] == -1 and ch[3, 2] == 1
    np.testing.assert_allclose(ch[1, 1], np.exp(2j * np.pi / 6), atol=1e-15)


def test_group_convention_square_matches_library_builder():
    import qmbcuda_amd.models as LM

    for n1, n2 in [(4, 4), (2, 4), (3, 2), (6, 6)]:
        geom = OM.square_geometry(n1, n2)
        ge, ch = OM.translation_group(geom)
        ge2, fl2, ch2 = LM.translation_group(LM.Lattice(n1, n2, 1))
        assert np.array_equal(ge, ge2)
        assert np.array_equal(ch, ch2)
        assert not fl2.any()
    geom = OM.triangular_geometry(3, 3)
    ge, ch = OM.translation_group(geom)
    ge2, _, ch2 = LM.translation_group(LM.Lattice(3, 3, 1))
    assert np.array_equal(ge, ge2) and np.array_equal(ch, ch2)


def test_term_table_matches_library_builder():
    import qmbcuda_amd.models as LM

    for n, J2, d in [(8, 0.0, 1.0), (10, 0.5, 0.7)]:
        geom, info = OM.chain_xxz_info(n, 1.0, d, J2)
        ref = OM.conjugated_term_table(OM.heisenberg_operator(geom, info))
        lat, inf2 = LM.chain(n, 1.0, d, J2)
        mine = LM.term_table(lat, inf2)
        for a, b in zip(ref[:4], mine[:4]):
            assert np.array_equal(a, b)
    geom, info = OM.square_xxz_info(4, 4, 1.0, 1.0)
    ref = OM.conjugated_term_table(OM.heisenberg_operator(geom, info))
    mine = LM.term_table(*LM.square(4, 4, 1.0, 1.0))
    for a, b in zip(ref[:4], mine[:4]):
        assert np.array_equal(a, b)


def test_extended_xxx_shells():
    # the active reference example: 2x4 square, {1, 0.15} (examples.cu:885-913)
    geom = OM.square_geometry(2, 4)
    info = OM.extended_xxx_info(geom, [1.0, 0.15])
    # nearest shell: |d| = 1 -> (+-1,0),(0,+-1); next: sqrt2 -> 4 diagonals
    nn = [r for r, J in zip(info.Jxy_terms, info.Jxy_couplings) if J == 1.0]
    nnn = [r for r, J in zip(info.Jxy_terms, info.Jxy_couplings) if J == 0.15]
    assert sorted(map(tuple, nn)) == sorted([(-1, 0, 0, 0), (1, 0, 0, 0), (0, -1, 0, 0), (0, 1, 0, 0)])
    assert len(nnn) == 4 and all(abs(r[0]) == 1 and abs(r[1]) == 1 for r in nnn)


@pytest.mark.parametrize("n,k,flip", [(8, 4, False), (10, 5, True), (9, 3, False), (12, 6, True)])
def test_orbits_match_brute_force(n, k, flip):
    case = chain_case(n, flip=flip)
    orb = oracle_orbits(case, k)
    brute = brute_orbits(n, k, case["group_el"], case["flip"])
    reps = sorted(brute)
    assert orb.n_orb == len(reps)
    assert orb.reps.tolist() == reps
    G = case["group_el"].shape[0]
    for o, r in enumerate(reps):
        stab, _ = brute[r]
        assert orb.stab[o] == len(stab)
        for a in range(orb.n_irreps):
            chi = case["chars"][a]
            compatible = all(abs(chi[g] - 1) < 1e-12 for g in stab)
            nsum = orb.norm_sum[o * orb.n_irreps + a]
            if compatible:
                assert abs(nsum - math.sqrt(G * len(stab)) / G) < 1e-12
            else:
                assert nsum < 1e-6


def test_orbit_index_map_is_canonical():
    case = chain_case(10, flip=True)
    orb = oracle_orbits(case, 5)
    G = case["group_el"].shape[0]
    for t, s in enumerate(orb.bas[:300]):
        o, g = orb.orbit_idx[2 * t], orb.orbit_idx[2 * t + 1]
        rep = int(orb.reps[o])
        img = oracle.apply_elem(rep, case["group_el"][g], case["flip"][g], 10)
        assert img == int(s)
        # smallest such g
        for h in range(g):
            assert oracle.apply_elem(rep, case["group_el"][h], case["flip"][h], 10) != int(s)


@pytest.mark.parametrize("label,case,n_up", [
    ("chain10_j2_xxz", chain_case(10, 1.0, 0.7, 0.3), 5),
    ("chain10_flip", chain_case(10, 1.0, 1.0, 0.45, flip=True), 5),
    ("chain9_odd", chain_case(9, 1.0, 1.3, 0.2), 4),
    ("square4x3", square_case(4, 3, 1.0, 0.8), 6),
    ("square3x3_flipless", square_case(3, 3, 1.0, 1.0), 4),
])
def test_spectral_completeness(label, case, n_up):
    """Union of the sector spectra (every irrep, null rows dropped) == dense spectrum."""
    dense = dense_spectrum(case, n_up)
    orb = oracle_orbits(case, n_up)
    parts = []
    for a in range(orb.n_irreps):
        ell = oracle_ell(case, orb, a)
        E = ell_dense_nonnull(ell, orb, a)
        if E.shape[0] == 0:
            continue
        np.testing.assert_allclose(E, E.conj().T, atol=1e-12)
        parts.append(np.linalg.eigvalsh(E))
    union = np.sort(np.concatenate(parts))
    assert len(union) == len(dense)
    np.testing.assert_allclose(union, dense, atol=1e-10)


def test_spmv_matches_dense():
    case = chain_case(12, 1.0, 0.9, 0.4)
    orb = oracle_orbits(case, 6)
    for a in (0, 3, 6):
        ell = oracle_ell(case, orb, a)
        rng = np.random.default_rng(a)
        x = rng.normal(size=orb.n_orb) + (1j * rng.normal(size=orb.n_orb) if ell.is_complex else 0)
        y = ell.spmv(x)
        np.testing.assert_allclose(y, ell.dense() @ x, atol=1e-12)


def test_lanczos_known_answers():
    # C1: N=16 chain, k=0 (known E0 from the survey's FP64 scipy run)
    case = chain_case(16)
    orb = oracle_orbits(case, 8)
    ell = oracle_ell(case, orb, 0)
    st, e0, it, a, b = ell.lanczos_converge(200, 2212456, tol=1e-13)
    assert st == 0
    assert abs(e0 - (-7.142296360617)) < 1e-11
    # Majumdar-Ghosh point: E0 = -3N/8 (k=0 flip-even for N/2 even)
    for n in (12, 16):
        case = chain_case(n, 1.0, 1.0, 0.5, flip=True)
        orb = oracle_orbits(case, n // 2)
        ell = oracle_ell(case, orb, 0)
        st, e0, it, a, b = ell.lanczos_converge(200, 7, tol=1e-13)
        assert st == 0
        assert abs(e0 + 3 * n / 8) < 1e-10


def test_small_closed_forms():
    # 4-ring Heisenberg: E0 = -2 (k=0 sector of the n_up=2 sector)
    case = chain_case(4)
    orb = oracle_orbits(case, 2)
    ell = oracle_ell(case, orb, 0)
    E = np.linalg.eigvalsh(ell_dense_nonnull(ell, orb, 0))
    assert abs(E[0] + 2.0) < 1e-13
    # 2 sites: a 2-site "chain" with +-1 offsets double counts the single bond
    # (each bond appears twice with periodic wrap); the dimer singlet is -3/4 per bond
    geom = OM.square_geometry(2, 1)
    info = OM.HeisenbergInfo(Jxy_terms=[[1, 0, 0, 0]], Jxy_couplings=[0.5], Jz_terms=[[1, 0, 0, 0]],
                             Jz_couplings=[0.5])
    terms = OM.heisenberg_operator(geom, info)
    H = OM.dense_sector_hamiltonian(2, 1, terms)
    assert abs(np.linalg.eigvalsh(H)[0] + 0.75) < 1e-14


def test_lanczos_matches_dense_complex_sector():
    case = square_case(4, 3, 1.0, 0.8)
    orb = oracle_orbits(case, 6)
    for a in (1, 4, 7):
        ell = oracle_ell(case, orb, a)
        assert ell.is_complex
        E = np.linalg.eigvalsh(ell_dense_nonnull(ell, orb, a))
        st, e0, it, al, be = ell.lanczos_converge(300, 11, tol=1e-13)
        assert st == 0
        assert abs(e0 - E[0]) < 1e-10 * max(1, abs(E[0]))


def test_eigvec_second_pass_residual():
    case = chain_case(14, 1.0, 1.0, 0.2)
    orb = oracle_orbits(case, 7)
    ell = oracle_ell(case, orb, 0)
    st, e0, it, a, b = ell.lanczos_converge(200, 99, tol=1e-13)
    assert st == 0
    w, z = oracle.tridiag_eig(a, b[1:len(a)])
    assert abs(w[0] - e0) < 1e-12
    res, x = ell.eigvec_residual(99, z[:, 0], w[0])
    # stopped before orthogonality is lost: a unit vector with a small residual
    assert abs(np.linalg.norm(x) - 1) < 1e-8
    assert res < 1e-5
    E = ell.dense()
    assert abs(x @ (E @ x) - e0) < 1e-10


def test_tridiag_ql_vs_numpy():
    rng = np.random.default_rng(3)
    for m in (1, 2, 5, 40, 150):
        d = rng.normal(size=m)
        e = rng.normal(size=m - 1)
        w, z = oracle.tridiag_eig(d, e)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(T), atol=1e-12)
        np.testing.assert_allclose(T @ z, z * w, atol=1e-10)


def _splitmix(seed, c):
    m = (1 << 64) - 1
    z = (seed + (c + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return ((z >> 11) + 1) * 2.0 ** -53


def test_start_vector_is_splitmix64():
    v = oracle.start_vector(2212456, 50)
    assert v.tolist() == [_splitmix(2212456, i) for i in range(50)]
    vc = oracle.start_vector(5, 10, True)
    assert vc.real.tolist() == [_splitmix(5, 2 * i) for i in range(10)]
    assert vc.imag.tolist() == [_splitmix(5, 2 * i + 1) for i in range(10)]
    assert np.all(v > 0) and np.all(v <= 1)


def test_planted_mistake_is_caught():
    """A wrong character sign (chi -> conj chi) must break spectral completeness
    only where it matters, and a dropped Sz factor must break the dense match."""
    case = chain_case(8, 1.0, 0.6, 0.0)
    orb = oracle_orbits(case, 4)
    dense = dense_spectrum(case, 4)
    bad = dict(case)
    op, site, sc, nops, hcol = case["table"]
    bad["table"] = (op, site, sc * np.where(op.reshape(-1, hcol)[:, 0] == 0, 0.5, 1.0), nops, hcol)
    parts = []
    for a in range(orb.n_irreps):
        ell = oracle_ell(bad, orb, a)
        E = ell_dense_nonnull(ell, orb, a)
        if E.shape[0]:
            parts.append(np.linalg.eigvalsh(E))
    assert not np.allclose(np.sort(np.concatenate(parts)), dense, atol=1e-6)


@pytest.mark.parametrize("label,case,n_up,alphas", [
    ("chain12_flip", chain_case(12, 1.0, 1.0, 0.3, flip=True), 6, [0, 5, 12, 17]),
    ("square4x4_flip", square_case(4, 4, 1.0, 0.8, flip=True), 8, [0, 5, 26]),
    ("chain13_odd", chain_case(13, 1.0, 0.7, 0.2), 6, [0, 2]),
])
def test_rows_from_reps_equal_full_table_build(label, case, n_up, alphas):
    """The sampled-row builder (brute-force orbit data, binary search in the
    representative list; used where the sector is too large for the full
    tables) equals the full-table build row for row, in both layouts."""
    orb = oracle_orbits(case, n_up)
    rng = np.random.default_rng(3)
    for a in alphas:
        ref = oracle_ell(case, orb, a)
        cplx = ref.is_complex
        rows = np.sort(rng.choice(orb.n_orb, min(orb.n_orb, 300), replace=False))
        ind, val, track = oracle.ell_rows_from_reps(case["n_sites"], case["group_el"], case["flip"], case["chars"],
                                                    case["table"], a, orb.reps, rows, ref.W, cplx)
        W = ref.W
        for i, r in enumerate(rows):
            n = int(ref.track[r])
            assert track[i] == n
            assert np.array_equal(ind[i * W:i * W + max(n, 1)], ref.ind[r * W:r * W + max(n, 1)])
            assert np.array_equal(np.asarray(val)[i * W:i * W + n], ref.values[r * W:r * W + n])
        # compact list: only the irrep's compatible orbits; columns are compact row indices
        keep = ~orb.null_mask(a)
        creps = orb.reps[keep]
        cidx = np.cumsum(keep) - 1
        crow = rng.choice(len(creps), min(len(creps), 200), replace=False)
        ind, val, track = oracle.ell_rows_from_reps(case["n_sites"], case["group_el"], case["flip"], case["chars"],
                                                    case["table"], a, creps, crow, W, cplx)
        full = np.flatnonzero(keep)[crow]
        for i, r in enumerate(full):
            n = int(ref.track[r])
            assert track[i] == n
            assert np.array_equal(ind[i * W + 1:i * W + n], cidx[ref.ind[r * W + 1:r * W + n]])
            assert np.array_equal(np.asarray(val)[i * W:i * W + n], ref.values[r * W:r * W + n])


def test_rows_from_reps_detects_missing_and_false_representatives():
    case = chain_case(12, 1.0, 1.0, 0.3, flip=True)
    orb = oracle_orbits(case, 6)
    W = OM.max_terms_rule(len(case["table"][3]), 6, 12)
    reps = orb.reps
    # a representative dropped from the list: some row's target orbit is then missing
    drop = np.delete(reps, len(reps) // 2)
    with pytest.raises(ValueError, match="missing"):
        oracle.ell_rows_from_reps(12, case["group_el"], case["flip"], case["chars"], case["table"], 0, drop,
                                  np.arange(len(drop)), W, False)
    # a listed state that is not its orbit's minimum
    bad = reps.copy()
    s = int(bad[5])
    img = max(oracle.apply_elem(s, case["group_el"][g], int(case["flip"][g]), 12) for g in range(len(case["flip"])))
    bad[5] = img
    with pytest.raises(ValueError, match="minimum"):
        oracle.ell_rows_from_reps(12, case["group_el"], case["flip"], case["chars"], case["table"], 0, bad,
                                  np.array([5]), W, False)


def test_burnside_count_matches_orbit_build():
    """The Burnside counting rule used for the full-size sector tests, against the
    oracle's orbit build (with and without spin flip, odd N, 2-D)."""
    for case, n_up in ((chain_case(12, flip=True), 6), (square_case(4, 4, flip=True), 8), (chain_case(13), 6),
                       (square_case(4, 3), 6), (chain_case(16, flip=True), 8)):
        orb = oracle.Orbits(case["n_sites"], n_up, case["group_el"], case["chars"][:1], flip=case["flip"])
        assert burnside_count(case, n_up) == orb.n_orb
    # the configured sector sizes
    assert burnside_count(chain_case(28, flip=True), 14) == 716728
    assert burnside_count(chain_case(30), 15) == 5170604


def test_locality_order_by_hand():
    """The locality slot order (terms sorted on lowest site, then highest site,
    descending, stable; operator-free terms first) on a table whose order is
    worked out by hand, and the same matrix from the reordered table: every row
    holds the same off-diagonal (column, value) pairs, only in another slot
    order, and the same diagonal up to the rounding of its sum (the diagonal
    terms are added in the new order)."""
    Z, P, M = OM.OP_SZ, OM.OP_SP, OM.OP_SM
    op = np.array([Z, Z, M, P, Z, Z, M, P, 0, 0, P, M, Z, 0], np.int32)
    site = np.array([0, 1, 1, 2, 3, 0, 2, 3, 0, 0, 3, 0, 2, 0], np.int32)
    scal = np.arange(1, 8).astype(np.complex128)
    n_ops = np.array([2, 2, 2, 2, 0, 2, 1], np.int32)
    # keys (lo, hi): (0,1) (1,2) (0,3) (2,3) none (0,3) (2,2)
    op2, site2, scal2, n_ops2, mo = oracle.locality_order((op, site, scal, n_ops, 2))
    assert np.array_equal(scal2.real.astype(int) - 1, [4, 3, 6, 1, 2, 5, 0])
    assert np.array_equal(n_ops2, [0, 2, 1, 2, 2, 2, 2]) and mo == 2
    assert np.array_equal(site2.reshape(-1, 2)[1], [2, 3]) and np.array_equal(op2.reshape(-1, 2)[1], [M, P])
    for case, n_up, alphas in ((chain_case(12, 1.0, 0.7, 0.3, flip=True), 6, (0, 5)),
                               (square_case(4, 3, 1.0, 0.8), 6, (1, 7))):
        orb = oracle_orbits(case, n_up)
        loc = dict(case, table=oracle.locality_order(case["table"]))
        for a in alphas:
            e1, e2 = oracle_ell(case, orb, a), oracle_ell(loc, orb, a)
            assert np.array_equal(e1.track, e2.track)
            W = e1.W
            key = lambda t: (t[0], complex(t[1]).real, complex(t[1]).imag)  # noqa: E731
            for r in range(e1.n_rows):
                n = int(e1.track[r])
                if n == 0:
                    continue
                assert e1.ind[r * W] == e2.ind[r * W] == r
                assert abs(e1.values[r * W] - e2.values[r * W]) <= 1e-14
                p1 = sorted(zip(e1.ind[r * W + 1:r * W + n].tolist(), e1.values[r * W + 1:r * W + n].tolist()), key=key)
                p2 = sorted(zip(e2.ind[r * W + 1:r * W + n].tolist(), e2.values[r * W + 1:r * W + n].tolist()), key=key)
                assert p1 == p2
            assert not np.array_equal(e1.ind, e2.ind)  # the slot order did change
