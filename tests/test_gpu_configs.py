"""GPU tests at the benchmark configurations' full sizes.

* 6x6 square Heisenberg, Sz=0, k=(0,0), flip-even (126,056,625 rows): the row
  count equals the Burnside count, the matrix is symmetric, sampled rows are
  bit-exact against the oracle's brute-force row builder, the ground-state
  energy per site agrees with the literature, and the learned start reaches the
  same eigenvalue in fewer steps.
* J1-J2 chain N=28 (the single-sector roofline workload): the whole basis and
  reference-layout ELL bit for bit against the oracle's full-table build.
* Learned start at its configured sizes: training on the N=20, 22, 24 chains
  over the 8x8 (Delta, J2/J1) grid, evaluation on N=30 held-out couplings.
"""
import sys
import time

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
import qmbcuda_amd as q
from oracle import model as OM
from qmbcuda_amd import models as LM
from tests.helpers import burnside_count, chain_case, oracle_ell, square_case

pytestmark = pytest.mark.gpu

# E0/N of the 6x6 Heisenberg antiferromagnet: Schulz, Ziman & Poilblanc,
# J. Phys. I France 6, 675 (1996), to its published digits
E0_PER_SITE_6X6 = -0.678872


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_c5_square6x6_full_size(ctx):
    say("C5 6x6 square k=0 flip-even: basis + ELL")
    case = square_case(6, 6, 1.0, 1.0, flip=True)
    lat, info = LM.square(6, 6, 1.0, 1.0)
    ge, fl, ch = LM.translation_group(lat, spin_flip=True)
    table = LM.term_table(lat, info)
    # the library's host tables are the oracle's
    assert np.array_equal(ge, case["group_el"]) and np.array_equal(fl, case["flip"])
    for a, b_ in zip(table, case["table"]):
        assert np.array_equal(np.asarray(a), np.asarray(b_))
    g = q.Group(ge, ch, fl)
    t = q.Terms(36, table)
    b = q.Basis(ctx, g, 18, 0, q.LAYOUT_COMPACT)
    n = b.n_rows
    assert n == burnside_count(case, 18) == 126056625
    # slots in locality order, as the benchmark builds the sector
    e = q.Ell(ctx, b, t, locality_slots=True)
    inf = e.info()
    assert inf.n_rows == n and inf.max_row_len <= 73 and inf.packed == 1

    say("C5 sampled rows vs oracle")
    reps, _, _ = b.export()
    assert np.all(reps[1:] > reps[:-1])
    rng = np.random.default_rng(2026)
    W = int(inf.max_terms)
    starts = np.sort(rng.integers(0, n - 100, 100))
    starts = np.concatenate([[0, n - 100], starts])
    rows = np.concatenate([np.arange(s, s + 100) for s in starts])
    ind_or, val_or, tr_or = oracle.ell_rows_from_reps(36, case["group_el"], case["flip"], case["chars"],
                                                      oracle.locality_order(case["table"]), 0, reps, rows, W,
                                                      False)
    for i, s in enumerate(starts):
        ind, val, tr = e.export_rows(int(s), 100)
        sl = slice(i * 100 * W, (i + 1) * 100 * W)
        assert np.array_equal(tr, tr_or[i * 100:(i + 1) * 100]), s
        assert np.array_equal(ind, ind_or[sl]), s
        assert np.array_equal(val, val_or[sl]), s
    nnz_sample = int(tr_or.astype(np.int64).sum())
    print(f"sampled {len(rows)} rows, {nnz_sample} entries bit-exact", flush=True)
    del reps

    say("C5 symmetry <x,Hy> = <Hx,y>; Hx on the sampled rows vs the oracle")
    x, y = rng.normal(size=n), rng.normal(size=n)
    hx = e.spmv(x)  # the streamed SpMV kernel (the matrix does not fit the MALL)
    lhs, rhs = x @ e.spmv(y), hx @ y
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)
    # the same products in the same slot order: bit for bit
    assert np.array_equal(hx[rows], oracle.spmv_rows(W, ind_or, val_or, tr_or, x))
    del x, y, hx

    say("C5 Lanczos, partial reorthogonalisation")
    ev, rep, _, _ = q.lanczos(e, max_iter=100, tol=1e-13, reorth=2, check_every=1)
    assert rep.converged
    print(f"E0 = {ev[0]:.12f}, E0/N = {ev[0] / 36:.10f}, {rep.iters} steps", flush=True)
    assert abs(ev[0] / 36 - E0_PER_SITE_6X6) < 5e-7

    say("C5 Lanczos, selective orthogonalisation (the benchmark's scheme)")
    ev_s, rep_s, _, _ = q.lanczos(e, max_iter=100, tol=1e-13, reorth=3, check_every=1)
    assert rep_s.converged and rep_s.n_reorth > 0
    assert abs(ev_s[0] - ev[0]) < 1e-10

    say("C5 ground-state observables")
    ev_x, rep_x, _, xg = q.lanczos(e, max_iter=100, tol=1e-13, reorth=2, check_every=1, want_eigvec=True)
    assert ev_x[0] == ev[0] and rep_x.converged
    nx = float(xg @ xg)
    eh, _ = b.expectation(t, xg)
    assert abs(eh.real / nx - ev[0]) < 1e-9 * abs(ev[0])
    # the singlet ground state is SU(2) and point-group invariant: every bond
    # carries E0/72 of S_i.S_j and a third of it in Sz_i Sz_j (to the accuracy
    # of the vector, ~sqrt(tol) for an operator that does not commute with H)
    for j in (1, 6):  # the horizontal and the vertical neighbour of site 0
        c, _ = b.expectation(q.Terms(36, LM.operator_table(LM.szsz(0, j))), xg)
        print(f"<Sz_0 Sz_{j}> = {c.real / nx:.12f} (E0/216 = {ev[0] / 216:.12f})", flush=True)
        assert abs(c.real / nx - ev[0] / 216) < 1e-6
    del xg

    say("C5 learned start")
    net = q.Mlp(ctx, seed=20261202)
    keep = []
    for (n1, n2) in ((4, 4), (4, 6), (6, 4)):
        for d in (0.8, 1.0, 1.2):
            lat_s, info_s = LM.square(n1, n2, 1.0, d)
            ge_s, fl_s, ch_s = LM.translation_group(lat_s, spin_flip=True)
            g_s = q.Group(ge_s, ch_s, fl_s)
            t_s = q.Terms(n1 * n2, LM.term_table(lat_s, info_s))
            b_s = q.Basis(ctx, g_s, n1 * n2 // 2, 0, q.LAYOUT_COMPACT)
            e_s = q.Ell(ctx, b_s, t_s)
            net.set_windows(b_s, LM.square_windows(lat_s), LM.sublattice_mask(lat_s))
            _, _, _, u = q.lanczos(e_s, max_iter=300, want_eigvec=True)
            net.add_sample(b_s, u, d, 0.0)
            keep += [g_s, t_s, b_s, e_s]
    net.train(300, 3e-3)
    net.set_windows(b, LM.square_windows(lat), LM.sublattice_mask(lat))
    ev_l, rep_l, fb, _ = net.lanczos(b, e, 1.0, 0.0, max_iter=100, tol=1e-13, reorth=2, check_every=1)
    print(f"learned: {rep_l.iters} steps (plain {rep.iters}), fallback {fb}", flush=True)
    assert rep_l.converged and not fb
    assert abs(ev_l[0] - ev[0]) < 1e-10 * abs(ev[0])
    assert rep_l.iters < rep.iters


def test_c2_full_size_bit_exact(ctx):
    """The J1-J2 chain N=28 (k=0, flip-even): every representative, stabiliser,
    column, value and row length of the reference-layout ELL against the oracle's
    full-table build (SetUpSRS over all 40,116,600 sector states)."""
    say("C2 N=28 oracle full-table build (CPU)")
    t0 = time.perf_counter()
    case = chain_case(28, 1.0, 1.0, 0.5, flip=True)
    orb = oracle.Orbits(28, 14, case["group_el"], case["chars"][:1], flip=case["flip"])
    ref = oracle_ell(case, orb, 0)
    print(f"oracle: {orb.n_orb} orbits, {time.perf_counter() - t0:.1f} s", flush=True)
    say("C2 N=28 GPU reference layout")
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(28, case["table"])
    b = q.Basis(ctx, g, 14, 0, q.LAYOUT_REFERENCE)
    reps, stab, nul = b.export()
    assert b.n_rows == orb.n_orb == 716728
    assert np.array_equal(reps, orb.reps)
    assert np.array_equal(stab.astype(np.int32), orb.stab)
    assert np.array_equal(nul.astype(bool), orb.null_mask(0))
    e = q.Ell(ctx, b, t, max_terms=ref.W)
    ind, val, track = e.export_reference()
    assert np.array_equal(track, ref.track)
    assert np.array_equal(ind, ref.ind)
    assert np.array_equal(val, ref.val)
    # the SpMV of the whole matrix, bit for bit
    x = np.random.default_rng(1).uniform(-1, 1, b.n_rows)
    assert np.array_equal(e.spmv(x), ref.spmv(x))


def test_c4_learned_start_configured_sizes(ctx):
    """Training set of the learned-start configuration (N=20, 22, 24 chains, k=0,
    8x8 grid of Delta in [0.5, 1.5] and J2/J1 in [0, 0.6]); two held-out N=30
    couplings: fewer Lanczos steps than the random start, same E0 to 1e-10."""
    say("C4 training data N=20,22,24 x 8x8 grid")
    net = q.Mlp(ctx, seed=20261202)
    keep = []
    deltas, j2s = np.linspace(0.5, 1.5, 8), np.linspace(0.0, 0.6, 8)
    for n in (20, 22, 24):
        lat = LM.Lattice(n)
        ge, fl, ch = LM.translation_group(lat)
        g = q.Group(ge, ch)
        b = q.Basis(ctx, g, n // 2, 0, q.LAYOUT_COMPACT)
        keep += [g, b]
        for d in deltas:
            for j2 in j2s:
                lat_, info = LM.chain(n, 1.0, float(d), float(j2))
                t = q.Terms(n, LM.term_table(lat_, info))
                e = q.Ell(ctx, b, t)
                _, _, _, u = q.lanczos(e, max_iter=600, want_eigvec=True)
                net.add_sample(b, u, float(d), float(j2))
                keep.append(t)
    say("C4 training")
    L0, _ = net.loss_grad(want_grad=False)
    net.train(400, 3e-3)
    L1, _ = net.loss_grad(want_grad=False)
    print(f"loss {L0:.4f} -> {L1:.4f}", flush=True)
    assert L1 < 0.5 * L0
    lat = LM.Lattice(30)
    ge, fl, ch = LM.translation_group(lat)
    g = q.Group(ge, ch)
    b = q.Basis(ctx, g, 15, 0, q.LAYOUT_COMPACT)
    assert b.n_rows == 5170604
    rng = np.random.default_rng(20261202)
    for case_i in range(2):
        d, j2 = float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.0, 0.6))
        say(f"C4 N=30 held-out Delta={d:.4f} J2={j2:.4f}")
        lat_, info = LM.chain(30, 1.0, d, j2)
        table = LM.term_table(lat_, info)
        e = q.Ell(ctx, b, q.Terms(30, table))
        if case_i == 0:
            # sampled rows of the N=30 matrix (ring orbit minima) and the streamed
            # SpMV's products on them, bit for bit against the oracle
            say("C4 N=30 sampled rows and SpMV vs oracle")
            reps, _, _ = b.export()
            W = int(e.info().max_terms)
            starts = np.concatenate([[0, b.n_rows - 64], np.sort(rng.integers(0, b.n_rows - 64, 60))])
            rows = np.concatenate([np.arange(s0, s0 + 64) for s0 in starts])
            ind_or, val_or, tr_or = oracle.ell_rows_from_reps(30, ge, None, ch, table, 0, reps, rows, W, False)
            for i, s0 in enumerate(starts):
                ind, val, tr = e.export_rows(int(s0), 64)
                sl = slice(i * 64 * W, (i + 1) * 64 * W)
                assert np.array_equal(tr, tr_or[i * 64:(i + 1) * 64]) and np.array_equal(ind, ind_or[sl])
                assert np.array_equal(val, val_or[sl])
            x = rng.normal(size=b.n_rows)
            assert np.array_equal(e.spmv(x)[rows], oracle.spmv_rows(W, ind_or, val_or, tr_or, x))
            del reps, x
        ev_r, rep_r, _, _ = q.lanczos(e, max_iter=1500, tol=1e-13)
        ev_l, rep_l, fb, _ = net.lanczos(b, e, d, j2, max_iter=1500, tol=1e-13)
        print(f"iterations random {rep_r.iters}, learned {rep_l.iters}", flush=True)
        assert rep_r.converged and rep_l.converged and not fb
        assert abs(ev_l[0] - ev_r[0]) < 1e-10 * abs(ev_r[0])
        assert rep_l.iters < rep_r.iters


def test_c3_full_size_sectors_vs_oracle(ctx):
    """Batched uneven sectors at full size (N=24 chain): the batched solve of all
    72 (k, n_up) sectors against the oracle's own converged Lanczos on three of
    them (real k=0 and k=12, complex k=5), lowest eigenvalue to 1e-10, and all
    four requested eigenvalues against ARPACK (scipy eigsh) on the oracle's
    matrix: a Krylov method finds each distinct eigenvalue once, so they are the
    four lowest distinct ones (the n_up=16, k=12 sector holds two exact
    doublets); on the same three the whole reference-layout ELL (columns,
    values, row lengths) and the SpMV bit for bit."""
    lat, info = LM.chain(24, 1.0, 1.0, 0.0)
    ge, fl, ch = LM.translation_group(lat)
    g = q.Group(ge, ch)
    t = q.Terms(24, LM.term_table(lat, info))
    say("C3 72 sectors batched")
    keep, ells = [], []
    for n_up in (12, 14, 16):
        for k in range(24):
            b = q.Basis(ctx, g, n_up, k, q.LAYOUT_COMPACT)
            keep.append(b)
            ells.append(q.Ell(ctx, b, t))
    ev, reps = q.lanczos_batched(ells, n_eigs=4, max_iter=400, tol=1e-13, reorth=2)
    assert all(r.converged for r in reps)
    case = chain_case(24)
    for n_up, k in ((12, 0), (14, 5), (16, 12)):
        say(f"C3 oracle sector n_up={n_up} k={k}")
        orb = oracle.Orbits(24, n_up, case["group_el"], case["chars"][[k]], flip=None)
        cplx = not np.all(case["chars"][k].imag == 0)
        W = OM.max_terms_rule(len(case["table"][3]), n_up, 24)
        ref = oracle.Ell(orb, case["table"], 0, W, cplx)
        bref = q.Basis(ctx, g, n_up, k, q.LAYOUT_REFERENCE)
        eref = q.Ell(ctx, bref, t, value_type=q.COMPLEX if cplx else q.REAL, max_terms=W)
        ind, val, track = eref.export_reference()
        assert np.array_equal(track, ref.track) and np.array_equal(ind, ref.ind)
        assert np.array_equal(val.view(np.float64), ref.val)
        x = np.random.default_rng(k).normal(size=orb.n_orb)
        if cplx:
            x = x + 1j * np.random.default_rng(k + 100).normal(size=orb.n_orb)
        assert np.array_equal(eref.spmv(x), ref.spmv(x))
        del eref, bref
        st, e0, it, _, _ = ref.lanczos_converge(600, 2212456, tol=1e-13, check_every=5)
        assert st == 0
        got = ev[[12, 14, 16].index(n_up) * 24 + k]
        print(f"n_up={n_up} k={k}: batched {got[0]:.12f}, oracle {e0:.12f}", flush=True)
        assert abs(got[0] - e0) < 1e-10 * abs(e0)
        keep = np.flatnonzero(~orb.null_mask(0))
        dt = np.complex128 if cplx else np.float64

        def mv(v, keep=keep, ref=ref, dt=dt, n=orb.n_orb):
            xx = np.zeros(n, dt)
            xx[keep] = v
            return ref.spmv(xx)[keep]

        op = spla.LinearOperator((len(keep), len(keep)), matvec=mv, dtype=dt)
        w = np.sort(spla.eigsh(op, k=8, which="SA", tol=1e-13, ncv=48)[0])
        distinct = [w[0]]
        for v in w[1:]:
            if v - distinct[-1] > 1e-8:
                distinct.append(v)
        print(f"  lowest 4 batched {np.round(got, 10)}, eigsh distinct {np.round(distinct[:4], 10)}", flush=True)
        np.testing.assert_allclose(got, distinct[:4], rtol=0, atol=1e-10 * abs(e0))
