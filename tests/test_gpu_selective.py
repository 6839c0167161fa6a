"""Selective orthogonalisation (reorth = 3): the single-sector solver keeps
semi-orthogonality by orthogonalising against the Ritz vectors that have
converged (B. N. Parlett and D. S. Scott, Math. Comp. 33 (1979) 217-238)
instead of against every kept vector when loss of orthogonality is estimated
(reorth = 2, H. D. Simon 1984).

Pins: the lowest eigenvalues against dense diagonalisation of the oracle's
matrix, run to tight tolerances where Ritz pairs at both ends of the spectrum
converge (so good vectors are formed and used); with many wanted eigenvalues
the scheme falls back to full reorthogonalisation and still matches dense;
the full N=28 sector against the partial scheme and the Majumdar-Ghosh E0;
the rebuilt eigenvector against its residual bound; the options it refuses.
"""
import sys

import numpy as np
import pytest

import qmbcuda_amd as q
from tests.helpers import chain_case, ell_dense_nonnull, oracle_ell, oracle_orbits, square_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def objects(ctx, case, n_up, alpha, value_type=None):
    g = q.Group(case["group_el"], case["chars"], case["flip"])
    t = q.Terms(case["n_sites"], case["table"])
    b = q.Basis(ctx, g, n_up, alpha, q.LAYOUT_COMPACT)
    e = q.Ell(ctx, b, t, value_type=value_type)
    return g, t, b, e


def distinct(values, tol=1e-8):
    out = [values[0]]
    for v in values[1:]:
        if v - out[-1] > tol:
            out.append(v)
    return np.array(out)


CASES = [
    ("chain16_k0", lambda: chain_case(16), 8, 0),
    ("chain14_j2_k3", lambda: chain_case(14, 1.0, 1.0, 0.3), 7, 3),
    ("square4x4_flip", lambda: square_case(4, 4, 1.0, 1.0, flip=True), 8, 0),
]


@pytest.mark.parametrize("label,mk,n_up,alpha", CASES)
def test_selective_lowest_four_equal_dense(ctx, label, mk, n_up, alpha):
    case = mk()
    orb = oracle_orbits(case, n_up)
    ref = oracle_ell(case, orb, alpha)
    dense = distinct(np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, alpha)))
    say(f"selective {label} irrep={alpha} dim={len(dense)}")
    *_, e = objects(ctx, case, n_up, alpha, value_type=q.COMPLEX if ref.is_complex else q.REAL)
    ev, rep, _, _ = q.lanczos(e, n_eigs=4, max_iter=600, tol=1e-15, reorth=3)
    assert rep.converged
    np.testing.assert_allclose(ev, dense[:4], rtol=0, atol=1e-10 * max(1.0, abs(dense[0])))
    # at this tolerance the extreme Ritz pairs converge: good vectors were formed
    assert rep.n_reorth > 0
    # the partial scheme gives the same eigenvalues
    ev2, rep2, _, _ = q.lanczos(e, n_eigs=4, max_iter=600, tol=1e-15, reorth=2)
    np.testing.assert_allclose(ev, ev2, rtol=0, atol=1e-10 * max(1.0, abs(dense[0])))


def test_selective_many_eigenvalues_fall_back_to_full(ctx):
    """Eight wanted eigenvalues: the watched low end (9 pairs) and high end (4)
    hold more good vectors than are kept, so the scheme orthogonalises against
    every kept vector from then on; the result still equals dense."""
    case = chain_case(16)
    orb = oracle_orbits(case, 8)
    ref = oracle_ell(case, orb, 0)
    dense = distinct(np.linalg.eigvalsh(ell_dense_nonnull(ref, orb, 0)))
    say("selective chain16 n_eigs=8")
    *_, e = objects(ctx, case, 8, 0)
    ev, rep, _, _ = q.lanczos(e, n_eigs=8, max_iter=800, tol=1e-14, reorth=3)
    assert rep.converged
    np.testing.assert_allclose(ev, dense[:8], rtol=0, atol=1e-10 * max(1.0, abs(dense[0])))
    assert rep.n_reorth > 0


def test_selective_full_size_n28(ctx):
    """The N=28 J1-J2 chain at the Majumdar-Ghosh point, flip-even k=0 (716,728
    rows): the selective scheme's E0 equals the partial scheme's and -10.5."""
    case = chain_case(28, 1.0, 1.0, 0.5, flip=True)
    say("selective N=28 MG")
    *_, e = objects(ctx, case, 14, 0)
    ev3, rep3, _, _ = q.lanczos(e, max_iter=300, tol=1e-13, reorth=3)
    ev2, rep2, _, _ = q.lanczos(e, max_iter=300, tol=1e-13, reorth=2)
    assert rep3.converged and rep2.converged
    assert abs(ev3[0] - ev2[0]) < 1e-10 * 10.5
    assert abs(ev3[0] + 10.5) < 1e-10 * 10.5


def test_selective_eigenvector(ctx):
    case = chain_case(14, 1.0, 1.0, 0.2)
    orb = oracle_orbits(case, 7)
    ref = oracle_ell(case, orb, 0)
    E = ell_dense_nonnull(ref, orb, 0)
    say("selective eigenvector chain14")
    *_, e = objects(ctx, case, 7, 0)
    ev, rep, _, x = q.lanczos(e, max_iter=300, tol=1e-13, reorth=3, want_eigvec=True)
    assert abs(np.linalg.norm(x) - 1) < 1e-8
    bound = 2.0 * np.sqrt(1e-13) * max(1.0, abs(ev[0]))
    assert rep.eig_residual <= bound
    assert np.linalg.norm(E @ x - ev[0] * x) <= bound


def test_selective_refuses_fixed_iter_checkpoints_and_batches(ctx, tmp_path):
    """Selective orthogonalisation tests every step on the host: it is refused
    with fixed_iter, with checkpoints and by the batched solver (whose steps run
    inside a replayed graph); the batched solver still takes reorth 2 after it."""
    case = chain_case(12)
    *_, e = objects(ctx, case, 6, 0)
    with pytest.raises(q.QMBError):
        q.lanczos(e, max_iter=20, reorth=3, fixed_iter=True)
    with pytest.raises(q.QMBError):
        q.lanczos_checkpointed(e, str(tmp_path / "lz.qlz"), max_iter=50, reorth=3, every=5)
    with pytest.raises(q.QMBError, match="reorth"):
        q.lanczos_batched([e, e], reorth=3)
    ev, reps = q.lanczos_batched([e, e], reorth=2)
    assert all(r.converged for r in reps) and abs(ev[0, 0] - ev[1, 0]) <= 1e-10 * abs(ev[0, 0])
