"""Point operations combined with translations (reflections of a chain, C4
rotations of a square torus): product groups whose one-dimensional
representations split the momentum sectors that the operation fixes.

Pins: the library's closure and representation checks agree with an
independent composition table; the union of the point-group sectors' spectra
equals the spectrum of the translation sector they split (oracle, dense); on
the GPU the bases and ELLs are bit-exact against the oracle and the energies
match dense diagonalisation.
"""
import sys

import numpy as np
import pytest

import oracle
import qmbcuda_amd as q
import qmbcuda_amd.models as LM
from oracle import model as OM
from tests.helpers import oracle_ell

SZ, SP, SM = OM.OP_SZ, OM.OP_SP, OM.OP_SM


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def chain_reflection_case(n=12, J2=0.4):
    lat, info = LM.chain(n, 1.0, 1.0, J2)
    T = LM.translation_group(lat, spin_flip=True)
    ge, fl, ch, pairs = LM.product_group(T, LM.reflection_group(lat))
    return dict(n_sites=n, lat=lat, table=LM.term_table(lat, info), base=T, group_el=ge, chars=ch, flip=fl,
                pairs=pairs, terms=None)


def square_c4_case(L=4):
    lat, info = LM.square(L, L, 1.0, 1.0)
    T = LM.translation_group(lat, spin_flip=True)
    ge, fl, ch, pairs = LM.product_group(T, LM.rotation_group(lat, 4))
    return dict(n_sites=L * L, lat=lat, table=LM.term_table(lat, info), base=T, group_el=ge, chars=ch, flip=fl,
                pairs=pairs, terms=None)


CASES = [("chain12_reflection_flip", chain_reflection_case, 6), ("square4x4_c4_flip", square_c4_case, 8)]


def sector_spectrum(case, group_el, chars, flip, n_up, alpha):
    orb = oracle.Orbits(case["n_sites"], n_up, group_el, chars, flip=flip)
    keep = ~orb.null_mask(alpha)
    if not keep.any():
        return np.zeros(0)
    W = OM.max_terms_rule(len(case["table"][3]), n_up, case["n_sites"]) + 4
    real = np.all(np.asarray(chars)[alpha].imag == 0)
    ell = oracle.Ell(orb, case["table"], alpha, W, not real)
    return np.linalg.eigvalsh(ell.dense()[np.ix_(keep, keep)])


@pytest.mark.parametrize("label,mk,n_up", CASES)
def test_library_marks_the_representations(label, mk, n_up):
    """Every pair (translation irrep, point irrep): the library's flag equals the
    composition-table check of the Python builder."""
    case = mk()
    geT, flT, chT = case["base"]
    P = LM.reflection_group(case["lat"]) if "chain" in label else LM.rotation_group(case["lat"], 4)
    geP, flP, chP = P
    GB = len(geP)
    ge, fl, _, pairs = LM.product_group(case["base"], P)
    rows = np.array([[chT[i][e // GB] * chP[j][e % GB] for e in range(len(ge))]
                     for i in range(chT.shape[0]) for j in range(chP.shape[0])])
    g = q.Group(ge, rows, fl)
    valid = g.valid_irreps()
    expect = np.zeros(len(rows), bool)
    for (i, j) in pairs:
        expect[i * chP.shape[0] + j] = True
    assert np.array_equal(valid, expect)
    assert q.Group(case["group_el"], case["chars"], case["flip"]).valid_irreps().all()


def test_non_closed_element_list_is_refused():
    lat, _ = LM.chain(8)
    ge, fl, ch = LM.translation_group(lat)
    with pytest.raises(q.QMBError):
        q.Group(np.concatenate([ge, LM.reflection_group(lat)[0][1:]]), np.ones((1, 9)), None)


@pytest.mark.parametrize("label,mk,n_up", CASES)
def test_point_group_sectors_split_the_momentum_sector(label, mk, n_up):
    """Union over the point-group irreps of the product sectors' spectra = the
    spectrum of the translation (x spin flip) sector they refine."""
    case = mk()
    geT, flT, chT = case["base"]
    by_t = {}
    for row, (i, j) in enumerate(case["pairs"]):
        by_t.setdefault(i, []).append(row)
    assert len(by_t) == 4  # two fixed momenta x two flip parities
    for i, rows in by_t.items():
        whole = sector_spectrum(case, geT, chT, flT, n_up, i)
        parts = np.sort(np.concatenate([sector_spectrum(case, case["group_el"], case["chars"], case["flip"], n_up, r)
                                        for r in rows]))
        assert len(parts) == len(whole) > 0
        np.testing.assert_allclose(parts, whole, atol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("label,mk,n_up", CASES)
def test_gpu_point_group_sectors(ctx, label, mk, n_up):
    case = mk()
    orb = oracle.Orbits(case["n_sites"], n_up, case["group_el"], case["chars"], flip=case["flip"])
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(case["n_sites"], case["table"])
    for a in range(len(case["pairs"])):
        say(f"point group {label} irrep={a}")
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        reps, stab, nul = b.export()
        assert np.array_equal(reps, orb.reps)
        assert np.array_equal(stab.astype(np.int32), orb.stab)
        assert np.array_equal(nul.astype(bool), orb.null_mask(a))
        ref = oracle_ell({"table": case["table"], "chars": case["chars"]}, orb, a,
                         max_terms=OM.max_terms_rule(len(case["table"][3]), n_up, case["n_sites"]) + 4)
        real = not ref.is_complex
        e = q.Ell(ctx, b, t, value_type=q.REAL if real else q.COMPLEX, max_terms=ref.W)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        keep = ~orb.null_mask(a)
        if not keep.any():
            continue
        dense = np.linalg.eigvalsh(ref.dense()[np.ix_(keep, keep)])
        bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
        ec = q.Ell(ctx, bc, t, value_type=q.REAL if real else q.COMPLEX, max_terms=ref.W)
        ev, rep, _, _ = q.lanczos(ec, max_iter=400, tol=1e-13)
        assert rep.converged
        assert abs(ev[0] - dense[0]) < 1e-10 * max(1.0, abs(dense[0]))


def affine_ring_case(n=10, u=3, delta=0.7):
    """Sites s -> u^k s + t (mod n), k = 0..3 (3 has order 4 mod 10), t = 0..n-1:
    the rotations times four cosets (k = 0..3). The k = 2 coset holds the site
    reversal (u^2 = -1 mod n); the k = 1, 3 cosets hold none, so their images
    take the generic permutation path of the ring x coset orbit minimum. The
    Hamiltonian couples every pair equally (XXZ, all to all), so every
    permutation leaves it invariant. Irreps: trivial and (-1)^k."""
    els, ks = [], []
    for k in range(4):
        m = pow(u, k, n)
        for t in range(n):
            els.append([(m * s + t) % n for s in range(n)])
            ks.append(k)
    chars = np.array([[1.0] * len(els), [(-1.0) ** k for k in ks]], dtype=np.complex128)
    ops, sites, scal = [], [], []
    for i in range(n):
        for j in range(i + 1, n):
            ops += [SM, SP]
            sites += [i, j]
            scal.append(0.5)
            ops += [SP, SM]
            sites += [i, j]
            scal.append(0.5)
            ops += [SZ, SZ]
            sites += [i, j]
            scal.append(delta)
    table = (np.array(ops, np.int32), np.array(sites, np.int32), np.array(scal, np.complex128),
             np.full(len(scal), 2, np.int32), 2)
    return dict(n_sites=n, group_el=np.array(els, np.uint32), chars=chars, flip=None, table=table)


def test_affine_ring_group_is_closed():
    case = affine_ring_case()
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    assert g.valid_irreps().all()


@pytest.mark.gpu
def test_gpu_affine_ring_cosets(ctx):
    """Ring rotations times generic coset representatives: bases and ELLs
    bit-exact against the oracle, energies against dense diagonalisation."""
    case = affine_ring_case()
    n_up = 5
    orb = oracle.Orbits(case["n_sites"], n_up, case["group_el"], case["chars"], flip=None)
    g = q.Group(case["group_el"], case["chars"], None)
    t = q.Terms(case["n_sites"], case["table"])
    for a in range(2):
        say(f"affine ring n=10 irrep={a}")
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        reps, stab, nul = b.export()
        assert np.array_equal(reps, orb.reps)
        assert np.array_equal(stab.astype(np.int32), orb.stab)
        assert np.array_equal(nul.astype(bool), orb.null_mask(a))
        W = OM.max_terms_rule(len(case["table"][3]), n_up, case["n_sites"]) + 4
        ref = oracle_ell({"table": case["table"], "chars": case["chars"]}, orb, a, max_terms=W)
        e = q.Ell(ctx, b, t, value_type=q.REAL, max_terms=ref.W)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        keep = ~orb.null_mask(a)
        if keep.any():
            dense = np.linalg.eigvalsh(ref.dense()[np.ix_(keep, keep)])
            bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
            ec = q.Ell(ctx, bc, t, value_type=q.REAL, max_terms=ref.W)
            ev, rep, _, _ = q.lanczos(ec, max_iter=300, tol=1e-13)
            assert rep.converged and abs(ev[0] - dense[0]) < 1e-10 * max(1.0, abs(dense[0]))


@pytest.mark.gpu
def test_gpu_chain24_reflection_full_sectors(ctx):
    """A full-size product group: J1-J2 chain of 24 sites, translations x spin
    flip x site reversal (96 elements), n_up = 12, all eight irreps of the two
    reflection-fixed momenta (about 28,000 orbit rows each). Per irrep: the
    basis, the whole reference-layout ELL and one SpMV bit for bit against the
    oracle, and E0 within 1e-10 of the oracle's converged Lanczos. Per
    translation sector: its GPU E0 equals the lower of its two reflection
    sectors' GPU E0s."""
    case = chain_reflection_case(24, 0.4)
    n, n_up = 24, 12
    orb = oracle.Orbits(n, n_up, case["group_el"], case["chars"], flip=case["flip"])
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(n, case["table"])
    W = OM.max_terms_rule(len(case["table"][3]), n_up, n) + 4
    e_point = {}
    for a, (i, j) in enumerate(case["pairs"]):
        say(f"chain24 reflection irrep={a} (translation irrep {i}, reflection {j})")
        b = q.Basis(ctx, g, n_up, a, q.LAYOUT_REFERENCE)
        reps, stab, nul = b.export()
        assert np.array_equal(reps, orb.reps)
        assert np.array_equal(stab.astype(np.int32), orb.stab)
        assert np.array_equal(nul.astype(bool), orb.null_mask(a))
        ref = oracle_ell({"table": case["table"], "chars": case["chars"]}, orb, a, max_terms=W)
        vt = q.COMPLEX if ref.is_complex else q.REAL
        e = q.Ell(ctx, b, t, value_type=vt, max_terms=ref.W)
        ind, val, track = e.export_reference()
        assert np.array_equal(track, ref.track)
        assert np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        x = np.random.default_rng(a + 70).normal(size=orb.n_orb)
        if ref.is_complex:
            x = x + 1j * np.random.default_rng(a + 170).normal(size=orb.n_orb)
        assert np.array_equal(e.spmv(x), ref.spmv(x))
        st, e0, _, _, _ = ref.lanczos_converge(600, 2212456, tol=1e-13, check_every=5)
        assert st == 0
        bc = q.Basis(ctx, g, n_up, a, q.LAYOUT_COMPACT)
        ec = q.Ell(ctx, bc, t, value_type=vt, max_terms=ref.W)
        ev, rep, _, _ = q.lanczos(ec, max_iter=600, tol=1e-13)
        assert rep.converged and abs(ev[0] - e0) < 1e-10 * max(1.0, abs(e0))
        e_point.setdefault(i, []).append(ev[0])
    geT, flT, chT = case["base"]
    gT = q.Group(geT, chT, flT)
    assert len(e_point) == 4
    for i, parts in e_point.items():
        say(f"chain24 translation irrep={i}")
        bt = q.Basis(ctx, gT, n_up, i, q.LAYOUT_COMPACT)
        real = bool(np.all(np.asarray(chT)[i].imag == 0))
        et = q.Ell(ctx, bt, t, value_type=q.REAL if real else q.COMPLEX, max_terms=W)
        ev, rep, _, _ = q.lanczos(et, max_iter=600, tol=1e-13)
        assert rep.converged and abs(ev[0] - min(parts)) < 1e-10 * max(1.0, abs(ev[0]))
