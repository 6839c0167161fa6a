"""Failure paths of the solver, forced from the host (every kernel exits
normally): a tolerance not reached within max_iter, a Krylov space that ends
at the start vector (beta = 0), a plan that cannot fit in device memory, and
each fallback of the learned start vector."""
import sys

import numpy as np
import pytest

import qmbcuda_amd as q
from qmbcuda_amd import models as LM
from tests.helpers import chain_case, ell_dense_nonnull, oracle_ell, oracle_orbits

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def chain_sector(ctx, n, delta=1.0, j2=0.0, flip=False):
    lat, info = LM.chain(n, 1.0, delta, j2)
    ge, fl, ch = LM.translation_group(lat, spin_flip=flip)
    g = q.Group(ge, ch, fl if flip else None)
    t = q.Terms(n, LM.term_table(lat, info))
    b = q.Basis(ctx, g, n // 2, 0, q.LAYOUT_COMPACT)
    return g, t, b, q.Ell(ctx, b, t)


def test_not_converged_within_max_iter(ctx):
    say("NOT_CONVERGED: max_iter 3 on the N=16 chain")
    keep = chain_sector(ctx, 16)
    e = keep[3]
    with pytest.raises(q.QMBError) as ei:
        q.lanczos(e, max_iter=3, tol=1e-13, check_every=1)
    assert ei.value.status == 4
    ev, rep, kr, _ = q.lanczos(e, max_iter=3, tol=1e-13, check_every=1, allow_unconverged=True)
    assert rep.converged == 0 and rep.breakdown == 0 and rep.iters == 3
    assert np.isfinite(ev[0]) and rep.residual[0] > 0.0
    # the same run with enough steps converges from the same seed
    ev2, rep2, _, _ = q.lanczos(e, max_iter=200, tol=1e-13)
    assert rep2.converged and abs(ev2[0] - (-7.142296360617)) < 1e-11


def test_breakdown_from_exact_eigenvector_start(ctx):
    """The start vector is an eigenvector of the sector matrix: b[1] vanishes,
    the run stops with the breakdown flag set and returns that eigenvalue."""
    say("breakdown: exact eigenvector start, N=12 chain")
    case = chain_case(12, 1.0, 1.0, 0.3)
    orb = oracle_orbits(case, 6)
    M = ell_dense_nonnull(oracle_ell(case, orb, 0), orb, 0)
    w, V = np.linalg.eigh(M)
    keep = chain_sector(ctx, 12, 1.0, 0.3)
    e = keep[3]
    assert keep[2].n_rows == M.shape[0]
    for j in (0, 2):
        ev, rep, _, _ = q.lanczos(e, max_iter=100, tol=1e-13, check_every=1, start_vec=V[:, j].real.copy())
        assert rep.breakdown == 1 and rep.converged == 1 and rep.iters == 1
        assert abs(ev[0] - w[j]) < 1e-12 * max(1.0, abs(w[j]))


def test_plan_beyond_device_memory_is_refused(ctx):
    """Full reorthogonalisation keeps every Lanczos vector: 100000 steps of the
    716,728-row sector need ~573 GB; the plan is refused up front with
    QMB_ERR_OOM and the context keeps working."""
    say("OOM: reorth=1, max_iter=100000 on the N=28 sector")
    keep = chain_sector(ctx, 28, 1.0, 0.5, flip=True)
    e = keep[3]
    with pytest.raises(q.QMBError) as ei:
        q.lanczos(e, max_iter=100000, reorth=1, tol=1e-13)
    assert ei.value.status == 3
    assert "plan needs" in str(ei.value)
    ev, rep, _, _ = q.lanczos(e, max_iter=500, tol=1e-13)
    assert rep.converged and abs(ev[0] + 10.5) < 1e-10 * 10.5


def _learned_case(ctx, n=16, delta=0.9, j2=0.2):
    keep = chain_sector(ctx, n, delta, j2)
    net = q.Mlp(ctx, seed=5)
    return keep, net


def test_learned_start_unusable_vector_falls_back(ctx):
    """NaN or infinite network parameters make the learned vector unusable (NaN,
    infinite, or zero amplitudes everywhere): the seeded random start is used
    instead (fallback 1) and the result equals the plain run from that seed bit
    for bit."""
    keep, net = _learned_case(ctx)
    b, e = keep[2], keep[3]
    ev_p, rep_p, _, _ = q.lanczos(e, max_iter=400, tol=1e-13, check_every=10)
    theta0 = net.params()
    for label, bad in (("NaN weights", lambda th: np.full_like(th, np.nan)),
                       ("infinite log-amplitude bias", lambda th: _set_bias(th, np.inf)),
                       ("log-amplitude bias -inf: every amplitude zero", lambda th: _set_bias(th, -np.inf))):
        say(f"learned fallback 1: {label}")
        net.set_params(bad(theta0.copy()))
        x, fb_v = net.start_vector(b, 0.9, 0.2)
        assert fb_v
        ev, rep, fb, _ = net.lanczos(b, e, 0.9, 0.2, max_iter=400, tol=1e-13, check_every=10)
        assert fb == 1
        assert ev[0] == ev_p[0] and rep.iters == rep_p.iters and rep.converged
    net.set_params(theta0)


def _set_bias(theta, v):
    # parameter layout: W1[128][10], b1[128], W2[128][128], b2[128], W3[2][128], b3[2]
    theta[-2] = v
    return theta


def test_learned_start_not_converged_falls_back_to_random(ctx):
    """A learned run that misses the tolerance within max_iter is judged by its
    residual: the random-start run with the same options is made and returned
    (fallback 2); here neither converges in 4 steps, so the result is the plain
    4-step run's, bit for bit, with NOT_CONVERGED."""
    say("learned fallback 2: max_iter 4")
    keep, net = _learned_case(ctx)
    b, e = keep[2], keep[3]
    ev_p, rep_p, _, _ = q.lanczos(e, max_iter=4, tol=1e-13, check_every=1, allow_unconverged=True)
    ev, rep, fb, _ = net.lanczos(b, e, 0.9, 0.2, max_iter=4, tol=1e-13, check_every=1, allow_unconverged=True)
    assert fb == 2
    assert rep.converged == 0 and rep.iters == rep_p.iters
    assert ev[0] == ev_p[0]
    with pytest.raises(q.QMBError) as ei:
        net.lanczos(b, e, 0.9, 0.2, max_iter=4, tol=1e-13, check_every=1)
    assert ei.value.status == 4


def test_learned_start_bad_network_still_reaches_e0(ctx):
    """A deliberately bad network (weights scaled far beyond training, so the
    learned vector is a spiky, wrong-signed function of the spins) still gives
    the plain run's E0 to 1e-10: the seeded noise in the start keeps the ground
    state in reach and the residual decides when the run is done."""
    say("learned start from a bad network")
    keep, net = _learned_case(ctx, n=18, delta=1.0, j2=0.3)
    b, e = keep[2], keep[3]
    rng = np.random.default_rng(9)
    theta = net.params()
    net.set_params((theta * 40.0 + rng.normal(0.0, 3.0, theta.shape)).astype(np.float32))
    ev_p, rep_p, _, _ = q.lanczos(e, max_iter=600, tol=1e-13, check_every=1)
    ev, rep, fb, _ = net.lanczos(b, e, 1.0, 0.3, max_iter=600, tol=1e-13, check_every=1)
    print(f"bad network: {rep.iters} steps (random start {rep_p.iters}), fallback {fb}", flush=True)
    assert rep.converged and fb in (0, 2)
    assert abs(ev[0] - ev_p[0]) < 1e-10 * abs(ev_p[0])
