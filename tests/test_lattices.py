"""General term tables: a lattice with two sites per cell (honeycomb) and complex
XXZ couplings, through the library's model tables, the oracle and the GPU.

Pins: the library's term and group tables equal the oracle's restatement of
the reference's Heisenberg constructor and Tnm_group on a two-sublattice
geometry; every irrep's sector spectrum together reproduces the dense
full-sector spectrum built from the operator algebra (complex hopping, complex
characters); on the GPU the ELL is bit-exact against the oracle and the
energies match dense diagonalisation.
"""
import math
import sys

import numpy as np
import pytest

import oracle
import qmbcuda_amd as q
import qmbcuda_amd.models as LM
from oracle import model as OM
from tests.helpers import dense_spectrum, ell_dense_nonnull, oracle_ell, oracle_orbits


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


# honeycomb: A at the cell origin, B at (a1 + a2)/3; B in cell (i, j) bonds to A in
# (i, j), (i+1, j), (i, j+1). Rows (di, dj, target sublattice, source sublattice).
AB_ROWS = [[0, 0, 1, 0], [-1, 0, 1, 0], [0, -1, 1, 0]]   # Sp_B Sm_A
BA_ROWS = [[0, 0, 0, 1], [1, 0, 0, 1], [0, 1, 0, 1]]     # Sp_A Sm_B


def honeycomb(n1, n2, J=1.0, delta=0.8, phi=0.37):
    """Both views of the honeycomb XXZ model with a complex hopping phase:
    (oracle geometry, oracle info, library lattice, library info)."""
    hop = J / 2 * complex(math.cos(phi), math.sin(phi))
    sq3 = math.sqrt(3.0) / 2.0
    geom = OM.Geometry([0.0, 0.5], [0.0, sq3 / 3.0], [[1.0, 0.5], [0.0, sq3]], n1, n2)
    oinfo = OM.HeisenbergInfo(Jxy_terms=AB_ROWS + BA_ROWS, Jxy_couplings=[hop] * 3 + [hop.conjugate()] * 3,
                              Jz_terms=AB_ROWS, Jz_couplings=[J * delta] * 3)
    lat = LM.Lattice(n1, n2, 2)
    linfo = LM.XXZInfo(Jxy_terms=AB_ROWS + BA_ROWS, Jxy_couplings=[hop] * 3 + [hop.conjugate()] * 3,
                       Jz_terms=AB_ROWS, Jz_couplings=[J * delta] * 3)
    return geom, oinfo, lat, linfo


def case_of(n1, n2, **kw):
    geom, oinfo, lat, linfo = honeycomb(n1, n2, **kw)
    terms = OM.heisenberg_operator(geom, oinfo)
    ge, ch = OM.translation_group(geom)
    return dict(n_sites=geom.n_sites, terms=terms, table=OM.conjugated_term_table(terms), group_el=ge, chars=ch,
                flip=None), lat, linfo


def test_library_tables_equal_the_oracle_on_two_sublattices():
    case, lat, linfo = case_of(2, 3)
    got = LM.term_table(lat, linfo)
    for u, v in zip(got[:4], case["table"][:4]):
        assert np.array_equal(np.asarray(u), np.asarray(v))
    assert got[4] == case["table"][4]
    ge, fl, ch = LM.translation_group(lat)
    assert np.array_equal(ge, case["group_el"])
    np.testing.assert_array_equal(ch, case["chars"])


@pytest.mark.parametrize("n1,n2,n_up", [(2, 3, 6), (3, 2, 5)])
def test_honeycomb_complex_hopping_spectral_completeness(n1, n2, n_up):
    case, _, _ = case_of(n1, n2)
    dense = dense_spectrum(case, n_up)
    orb = oracle_orbits(case, n_up)
    parts = []
    for a in range(orb.n_irreps):
        E = ell_dense_nonnull(oracle_ell(case, orb, a, is_complex=True), orb, a)
        if E.shape[0]:
            np.testing.assert_allclose(E, E.conj().T, atol=1e-12)
            parts.append(np.linalg.eigvalsh(E))
    union = np.sort(np.concatenate(parts))
    assert len(union) == len(dense)
    np.testing.assert_allclose(union, dense, atol=1e-10)


@pytest.mark.gpu
def test_gpu_honeycomb_complex_hopping(ctx):
    case, lat, linfo = case_of(2, 4)
    n_up = 8
    orb = oracle_orbits(case, n_up)
    g = q.Group(case["group_el"], case["chars"])
    t = q.Terms(case["n_sites"], LM.term_table(lat, linfo))
    for a in range(orb.n_irreps):
        say(f"honeycomb 2x4 irrep={a}")
        ref = oracle_ell(case, orb, a, is_complex=True)
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        e = q.Ell(ctx, b, t, value_type=q.COMPLEX, max_terms=ref.W)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        rng = np.random.default_rng(a)
        x = rng.normal(size=orb.n_orb) + 1j * rng.normal(size=orb.n_orb)
        assert np.array_equal(e.spmv(x), ref.spmv(x))
        keep = ~orb.null_mask(a)
        if not keep.any():
            continue
        e0 = np.linalg.eigvalsh(ref.dense()[np.ix_(keep, keep)])[0]
        bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
        ec = q.Ell(ctx, bc, t, value_type=q.COMPLEX, max_terms=ref.W)
        ev, rep, _, _ = q.lanczos(ec, max_iter=300, tol=1e-13)
        assert rep.converged and abs(ev[0] - e0) < 1e-10 * max(1.0, abs(e0))


@pytest.mark.gpu
def test_gpu_honeycomb_3x4_full_sectors(ctx):
    """A larger two-sublattice case: honeycomb 3x4 (24 sites, C(24,12) states,
    225,440 orbit rows per irrep), complex hopping, three irreps: the whole
    reference-layout ELL and the SpMV bit for bit against the oracle, and E0
    within 1e-10 of the oracle's converged Lanczos."""
    case, lat, linfo = case_of(3, 4)
    n_up = 12
    orb = oracle_orbits(case, n_up)
    g = q.Group(case["group_el"], case["chars"])
    t = q.Terms(case["n_sites"], LM.term_table(lat, linfo))
    for a in (0, 1, 7):
        say(f"honeycomb 3x4 irrep={a}")
        ref = oracle_ell(case, orb, a, is_complex=True)
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        e = q.Ell(ctx, b, t, value_type=q.COMPLEX, max_terms=ref.W)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        rng = np.random.default_rng(a + 40)
        x = rng.normal(size=orb.n_orb) + 1j * rng.normal(size=orb.n_orb)
        assert np.array_equal(e.spmv(x), ref.spmv(x))
        st, e0, _, _, _ = ref.lanczos_converge(600, 2212456, tol=1e-13, check_every=5)
        assert st == 0
        bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
        ec = q.Ell(ctx, bc, t, value_type=q.COMPLEX, max_terms=ref.W)
        ev, rep, _, _ = q.lanczos(ec, max_iter=600, tol=1e-13)
        assert rep.converged and abs(ev[0] - e0) < 1e-10 * max(1.0, abs(e0))
