"""Batched multi-sector Lanczos (new pattern; no reference counterpart).

Pinned by equality with the sequential single-sector solves (same seed, same
recurrence) and, at small N, with dense diagonalisation of the oracle's sector
matrices."""
import numpy as np
import pytest

import qmbcuda_amd as q
from qmbcuda_amd import models as LM
from tests.helpers import chain_case, ell_dense_nonnull, oracle_ell, oracle_orbits

pytestmark = pytest.mark.gpu


def sector_ells(ctx, n, n_ups, J2=0.0, delta=1.0, locality_slots=False):
    lat, info = LM.chain(n, 1.0, delta, J2)
    ge, fl, ch = LM.translation_group(lat)
    g = q.Group(ge, ch)
    t = q.Terms(n, LM.term_table(lat, info))
    keep, ells, labels = [g, t], [], []
    for n_up in n_ups:
        for k in range(n):
            b = q.Basis(ctx, g, n_up, k, q.LAYOUT_COMPACT)
            if b.n_rows == 0:
                continue
            e = q.Ell(ctx, b, t, locality_slots=locality_slots)
            keep.append(b)
            ells.append(e)
            labels.append((n_up, k))
    return keep, ells, labels


def test_batched_matches_dense_small(ctx):
    print("[gpu-case] batched chain12 lowest-4", flush=True)
    keep, ells, labels = sector_ells(ctx, 12, [6, 4], J2=0.3, delta=0.9)
    ev, reps = q.lanczos_batched(ells, n_eigs=4, max_iter=300, tol=1e-13)
    case = chain_case(12, 1.0, 0.9, 0.3)
    for n_up in (6, 4):
        orb = oracle_orbits(case, n_up)
        for a in range(12):
            if (n_up, a) not in labels:
                continue
            i = labels.index((n_up, a))
            E = np.linalg.eigvalsh(ell_dense_nonnull(oracle_ell(case, orb, a), orb, a))
            want = E[:4]
            got = ev[i][: len(want)]
            assert reps[i].converged
            np.testing.assert_allclose(got, want, atol=1e-10 * max(1.0, abs(want[0])))


def test_batched_partial_reorthogonalisation_matches_dense(ctx):
    """reorth = 2 in the batch: the per-instance device decisions keep the lowest
    four eigenvalues exact while orthogonalising on only part of the steps."""
    print("[gpu-case] batched chain12 lowest-4 partial reorthogonalisation", flush=True)
    keep, ells, labels = sector_ells(ctx, 12, [6, 4], J2=0.3, delta=0.9)
    ev, reps = q.lanczos_batched(ells, n_eigs=4, max_iter=300, tol=1e-13, reorth=2)
    full, _ = q.lanczos_batched(ells, n_eigs=4, max_iter=300, tol=1e-13, reorth=1)
    case = chain_case(12, 1.0, 0.9, 0.3)
    partial_steps = 0
    for n_up in (6, 4):
        orb = oracle_orbits(case, n_up)
        for a in range(12):
            if (n_up, a) not in labels:
                continue
            i = labels.index((n_up, a))
            E = np.linalg.eigvalsh(ell_dense_nonnull(oracle_ell(case, orb, a), orb, a))
            want = E[:4]
            assert reps[i].converged
            np.testing.assert_allclose(ev[i][: len(want)], want, atol=1e-10 * max(1.0, abs(want[0])))
            np.testing.assert_allclose(ev[i][: len(want)], full[i][: len(want)], atol=1e-10 * max(1.0, abs(want[0])))
            assert reps[i].n_reorth <= reps[i].iters + 10
            partial_steps += reps[i].n_reorth < reps[i].iters
    assert partial_steps > 0


def test_batched_equals_sequential_ground_states(ctx):
    print("[gpu-case] batched chain16 ground states", flush=True)
    keep, ells, labels = sector_ells(ctx, 16, [8, 6])
    ev, reps = q.lanczos_batched(ells, n_eigs=1, max_iter=400, tol=1e-13)
    for i, e in enumerate(ells):
        seq, rep, _, _ = q.lanczos(e, max_iter=400, tol=1e-13)
        assert abs(ev[i][0] - seq[0]) < 1e-10 * max(1.0, abs(seq[0])), labels[i]
    # the N=16 chain ground state lives in (n_up=8, k=0)
    assert abs(ev[labels.index((8, 0))][0] - (-7.142296360617)) < 1e-10


@pytest.mark.parametrize("reorth", [1, 2])
def test_batched_vectors_are_ground_states(ctx, reorth):
    """qmb_lanczos_batched_vectors: every instance's vector (real and complex
    sectors) is the dense ground state of its oracle sector matrix up to a phase
    (|<x, g>| = 1 to 1e-9) with the convergence test's residual bound, and
    matches the single solve's want_eigvec vector; reorth = 0 is refused."""
    print(f"[gpu-case] batched vectors chain12 reorth={reorth}", flush=True)
    keep, ells, labels = sector_ells(ctx, 12, [6, 4], J2=0.3, delta=0.9)
    ev, reps, vecs = q.lanczos_batched(ells, max_iter=300, tol=1e-13, reorth=reorth, want_eigvec=True)
    case = chain_case(12, 1.0, 0.9, 0.3)
    orbs = {n_up: oracle_orbits(case, n_up) for n_up in (6, 4)}
    for i, (n_up, a) in enumerate(labels):
        x = vecs[i]
        assert reps[i].converged
        H = ell_dense_nonnull(oracle_ell(case, orbs[n_up], a), orbs[n_up], a)
        E, V = np.linalg.eigh(H)
        E0 = ev[i][0]
        assert abs(E0 - E[0]) < 1e-10 * max(1.0, abs(E[0]))
        nx = np.linalg.norm(x)
        assert abs(nx - 1.0) < 1e-6
        res = np.linalg.norm(H @ x - E0 * x)
        assert res < 2 * np.sqrt(1e-13) * max(1.0, abs(E0)), (labels[i], res)
        if E[1] - E[0] > 1e-6:  # non-degenerate ground state: the vector itself
            assert abs(abs(np.vdot(V[:, 0], x)) - nx) < 1e-9, labels[i]
        if reorth == 2 and i % 5 == 0:
            _, _, _, xs = q.lanczos(ells[i], max_iter=300, tol=1e-13, reorth=2, want_eigvec=True)
            assert abs(abs(np.vdot(xs, x)) - np.linalg.norm(xs) * nx) < 1e-9, labels[i]
    if reorth == 1:
        with pytest.raises(q.QMBError):
            q.lanczos_batched(ells[:2], max_iter=300, tol=1e-13, reorth=0, want_eigvec=True)


def test_batched_locality_slots(ctx):
    """The batch of locality-ordered matrices (the benchmark's matrices) gives the
    table-order batch's lowest four eigenvalues of every sector (1e-10)."""
    print("[gpu-case] batched chain12 locality slots", flush=True)
    keep, ells, labels = sector_ells(ctx, 12, [6, 4], J2=0.3, delta=0.9)
    keep2, ells2, labels2 = sector_ells(ctx, 12, [6, 4], J2=0.3, delta=0.9, locality_slots=True)
    assert labels == labels2
    ev, reps = q.lanczos_batched(ells, n_eigs=4, max_iter=300, tol=1e-13, reorth=2)
    ev2, reps2 = q.lanczos_batched(ells2, n_eigs=4, max_iter=300, tol=1e-13, reorth=2)
    for i in range(len(ells)):
        assert reps[i].converged and reps2[i].converged
        np.testing.assert_allclose(ev2[i], ev[i], atol=1e-10 * max(1.0, abs(ev[i][0])))


def test_batched_c3_full_size(ctx):
    """C3: N=24 chain, 24 momenta x n_up in {12, 14, 16}: 72 uneven sectors, lowest 4 each."""
    print("[gpu-case] batched C3 N=24 72 sectors", flush=True)
    keep, ells, labels = sector_ells(ctx, 24, [12, 14, 16])
    assert len(ells) == 72
    dims = [e.info().n_rows for e in ells]
    assert max(dims) == 112720 and min(dims) >= 30624
    ev, reps = q.lanczos_batched(ells, n_eigs=4, max_iter=400, tol=1e-13)
    assert all(r.converged for r in reps)
    # sequential check on a spread of sectors (real and complex)
    for i in (0, 1, 7, 12, 30, 49, 71):
        seq, rep, _, _ = q.lanczos(ells[i], n_eigs=4, max_iter=400, tol=1e-13)
        np.testing.assert_allclose(ev[i], seq, atol=1e-10 * max(1.0, abs(seq[0])))
    # SU(2): the lowest level of Sz=2 lies above the Sz=0 ground state
    e0 = min(ev[i][0] for i, (n_up, k) in enumerate(labels) if n_up == 12)
    e2 = min(ev[i][0] for i, (n_up, k) in enumerate(labels) if n_up == 14)
    assert e0 < e2
