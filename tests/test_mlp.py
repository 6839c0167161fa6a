"""Learned start vector: oracle pins (CPU) and GPU MFMA forward/backward checks.

CPU: the FP64 oracle's back-propagation equals central finite differences of
its own loss; its bfloat16 rounding equals torch's; the Marshall prior is the
exact sign of the unfrustrated chain's ground state (and a uniform vector is
orthogonal to it, which is why the prior is there); the start-vector noise
stream equals the C oracle's random start stream.
GPU: parameter initialisation bit-equal to the oracle's; pattern-table forward
pass and loss within bf16 tolerance of the oracle; the MFMA gradient against
the oracle gradient (same bf16 rounding points); the learned start vector
against the oracle's; training lowers the loss; a trained network starts
Lanczos closer to the ground state than the random start (fewer iterations,
same E0) on held-out couplings and a larger ring.
"""
import numpy as np
import pytest

from oracle import mlp as OMLP


def small_problem(n=10, delta=0.8, j2=0.3, seed=3):
    """A random target vector over the k=0 ring representatives (CPU only)."""
    from tests.helpers import chain_case, oracle_orbits

    case = chain_case(n)
    orb = oracle_orbits(case, n // 2)
    reps = orb.reps
    rng = np.random.default_rng(seed)
    u = rng.normal(size=len(reps))
    u /= np.linalg.norm(u)
    return reps, orb.stab, n, u, delta, j2


def test_oracle_backprop_matches_finite_differences():
    reps, stab, n, u, d, j2 = small_problem()
    pats = OMLP.window_patterns(reps, n)
    prior = OMLP.marshall_phase(reps, n)
    theta = OMLP.init_params(11) * 3.0
    L, g = OMLP.fidelity_loss(theta, pats, stab, n, u, d, j2, with_grad=True, round_like_gpu=False, prior=prior)
    rng = np.random.default_rng(0)
    idx = rng.choice(OMLP.N_PARAMS, 40, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (OMLP.fidelity_loss(tp, pats, stab, n, u, d, j2, round_like_gpu=False, prior=prior)
              - OMLP.fidelity_loss(tm, pats, stab, n, u, d, j2, round_like_gpu=False, prior=prior)) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_marshall_prior_is_the_unfrustrated_ground_state_sign():
    """Heisenberg ring N=12 (N/2 even), k=0: sign(u_r) = +-(-1)^(up spins on even
    sites) for every representative (Marshall's theorem), and the uniform vector
    (the fully polarised multiplet) has zero overlap with the singlet."""
    from tests.helpers import chain_case, oracle_ell, oracle_orbits

    n = 12
    case = chain_case(n, 1.0, 1.0, 0.0)
    orb = oracle_orbits(case, n // 2)
    H = oracle_ell(case, orb, 0).dense()
    w, V = np.linalg.eigh(H)
    u = V[:, 0]
    reps = orb.bas[orb.rep_rank]
    sgn = np.cos(OMLP.marshall_phase(reps, n))
    assert np.all(np.abs(u) > 1e-12)
    prod = np.sign(u) * sgn
    assert np.all(prod == prod[0])
    uniform = np.sqrt(n / orb.stab.astype(np.float64))
    assert abs(uniform @ u) < 1e-10 * np.linalg.norm(uniform)


def test_marshall_prior_on_the_square_lattice():
    """4x4 Heisenberg, k=(0,0), flip-even: the sublattice-A Marshall sign is the
    exact sign of every ground-state coefficient (bipartite lattice)."""
    from qmbcuda_amd import models as LM
    from tests.helpers import oracle_ell, oracle_orbits, square_case

    case = square_case(4, 4, 1.0, 1.0, flip=True)
    orb = oracle_orbits(case, 8)
    H = oracle_ell(case, orb, 0).dense()
    w, V = np.linalg.eigh(H)
    u = V[:, 0]
    lat, _ = LM.square(4, 4, 1.0, 1.0)
    sgn = np.cos(OMLP.marshall_phase(orb.bas[orb.rep_rank], 16, LM.sublattice_mask(lat)))
    prod = np.sign(u) * sgn
    assert np.all(np.abs(u) > 1e-12) and np.all(prod == prod[0])
    # the ring default (even sites) is not the square lattice's sublattice
    assert not np.all(np.sign(u) * np.cos(OMLP.marshall_phase(orb.bas[orb.rep_rank], 16)) == prod[0])


def test_square_windows():
    from qmbcuda_amd import models as LM

    lat, _ = LM.square(4, 6, 1.0, 1.0)
    W = LM.square_windows(lat)
    assert W.shape == (48, 8)
    for w in W:
        assert len(set(w.tolist())) == 8
    # window 0: rows 0-1, columns 0-3; window 24: rows 0-3, columns 0-1
    assert W[0].tolist() == [0, 1, 2, 3, 6, 7, 8, 9]
    assert W[24].tolist() == [0, 1, 6, 7, 12, 13, 18, 19]
    # every site is in 8 windows of each shape
    assert np.all(np.bincount(W[:24].ravel(), minlength=24) == 8)
    assert np.all(np.bincount(W[24:].ravel(), minlength=24) == 8)


def test_start_noise_is_the_random_start_stream():
    import oracle

    for seed in (0, 2212456, 2**63 + 5):
        want = 2.0 * oracle.start_vector(seed, 300) - 1.0
        assert np.array_equal(OMLP.start_noise(300, seed), want)


def test_bf16_rounding_matches_torch():
    import torch

    x = np.random.default_rng(1).normal(size=10000) * 10.0 ** np.random.default_rng(2).integers(-6, 6, 10000)
    want = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OMLP.bf16(x), want)


@pytest.mark.gpu
def test_mlp_init_forward_loss_and_gradient(ctx):
    import qmbcuda_amd as q
    from qmbcuda_amd import models as LM

    print("[gpu-case] mlp init/forward/grad", flush=True)
    net = q.Mlp(ctx, seed=20261202)
    th = net.params().astype(np.float64)
    assert np.array_equal(th, OMLP.init_params(20261202))
    theta = th * 3.0  # larger weights so tanh saturation and phases matter
    net.set_params(theta.astype(np.float32))
    theta = theta.astype(np.float32).astype(np.float64)
    # forward pass over all 256 patterns of two couplings
    out = net.forward([0.7, 1.3], [0.1, 0.55])
    for i, (d, j2) in enumerate([(0.7, 0.1), (1.3, 0.55)]):
        ref, _ = OMLP.forward(theta, OMLP.pattern_inputs(d, j2))
        err = np.abs(out[i] - ref)
        assert err.max() < 3e-2 and np.median(err) < 2e-3, err.max()
    # one sample: the exact ground state of an N=12 chain sector (k=0)
    lat, info = LM.chain(12, 1.0, 0.8, 0.3)
    ge, fl, ch = LM.translation_group(lat)
    g = q.Group(ge, ch)
    t = q.Terms(12, LM.term_table(lat, info))
    b = q.Basis(ctx, g, 6, 0, q.LAYOUT_COMPACT)
    e = q.Ell(ctx, b, t)
    ev, rep, _, u = q.lanczos(e, max_iter=300, want_eigvec=True)
    net.add_sample(b, u, 0.8, 0.3)
    L, grad = net.loss_grad()
    reps, stab, _ = b.export()
    pats = OMLP.window_patterns(reps, 12)
    prior = OMLP.marshall_phase(reps, 12)
    L_or, g_or = OMLP.fidelity_loss(theta, pats, stab, 12, u, 0.8, 0.3, with_grad=True, round_like_gpu=True,
                                    prior=prior)
    assert abs(L - L_or) < 2e-3 * max(1.0, abs(L_or))
    cos = grad @ g_or / (np.linalg.norm(grad) * np.linalg.norm(g_or))
    assert cos > 0.999, cos
    assert abs(np.linalg.norm(grad) / np.linalg.norm(g_or) - 1) < 3e-2
    # the big blocks separately: W2 (MFMA weight gradient) and W1 (through the MFMA dH1)
    for name, sl in (("W2", slice(1408, 1408 + 16384)), ("W1", slice(0, 1280))):
        a, r = grad[sl], g_or[sl]
        assert a @ r / (np.linalg.norm(a) * np.linalg.norm(r)) > 0.995, name
    # the learned start vector (before the noise admixture) against the oracle's
    x, fb = net.start_vector(b, 0.8, 0.3)
    assert not fb
    x_or = OMLP.start_vector(theta, pats, prior, stab, 12, 0.8, 0.3)
    assert x @ x_or / (np.linalg.norm(x) * np.linalg.norm(x_or)) > 0.999


@pytest.mark.gpu
def test_mlp_training_and_learned_start_gain(ctx):
    """Train on N=12,14,16 k=0 ground states over a 3x3 coupling grid; evaluate
    on N=20 at held-out couplings: the learned start needs fewer Lanczos steps
    than the seeded random start and reaches the same E0."""
    import qmbcuda_amd as q
    from qmbcuda_amd import models as LM

    print("[gpu-case] mlp training + gain", flush=True)
    keep = []
    net = q.Mlp(ctx, seed=7)

    def sector(n, d, j2):
        lat, info = LM.chain(n, 1.0, d, j2)
        ge, fl, ch = LM.translation_group(lat)
        g = q.Group(ge, ch)
        t = q.Terms(n, LM.term_table(lat, info))
        b = q.Basis(ctx, g, n // 2, 0, q.LAYOUT_COMPACT)
        e = q.Ell(ctx, b, t)
        keep.extend([g, t, b, e])
        return b, e

    for n in (12, 14, 16):
        for d in np.linspace(0.6, 1.4, 3):
            for j2 in np.linspace(0.0, 0.5, 3):
                b, e = sector(n, d, j2)
                ev, rep, _, u = q.lanczos(e, max_iter=400, want_eigvec=True)
                net.add_sample(b, u, float(d), float(j2))
    L0, _ = net.loss_grad(want_grad=False)
    hist = net.train(300, 3e-3)
    L1, _ = net.loss_grad(want_grad=False)
    print(f"loss {L0:.4f} -> {L1:.4f}", flush=True)
    assert L1 < 0.5 * L0
    gains = []
    for d, j2 in [(0.9, 0.2), (1.1, 0.35)]:
        b, e = sector(20, d, j2)
        ev_r, rep_r, _, _ = q.lanczos(e, max_iter=600, tol=1e-13, check_every=1)
        ev_l, rep_l, fb, _ = net.lanczos(b, e, d, j2, max_iter=600, tol=1e-13, check_every=1)
        print(f"N=20 ({d}, {j2}): iterations random {rep_r.iters}, learned {rep_l.iters}", flush=True)
        assert not fb and rep_r.converged and rep_l.converged
        assert abs(ev_l[0] - ev_r[0]) < 1e-10 * abs(ev_r[0])
        gains.append(rep_r.iters - rep_l.iters)
    assert min(gains) > 0, gains


@pytest.mark.gpu
def test_mlp_square_lattice_windows(ctx):
    """2-D windows: the GPU loss and gradient on a 4x4 k=0 flip-even sample
    against the oracle; after training on 4x4 and 4x6 ground states, the learned
    start needs fewer Lanczos steps on a 4x6 coupling it was not trained on."""
    import qmbcuda_amd as q
    from qmbcuda_amd import models as LM

    print("[gpu-case] mlp square-lattice windows", flush=True)
    keep = []

    def sector(n1, n2, d):
        lat, info = LM.square(n1, n2, 1.0, d)
        ge, fl, ch = LM.translation_group(lat, spin_flip=True)
        g = q.Group(ge, ch, fl)
        t = q.Terms(n1 * n2, LM.term_table(lat, info))
        b = q.Basis(ctx, g, n1 * n2 // 2, 0, q.LAYOUT_COMPACT)
        e = q.Ell(ctx, b, t)
        keep.extend([g, t, b, e])
        return lat, b, e

    net = q.Mlp(ctx, seed=11)
    lat, b, e = sector(4, 4, 1.0)
    W, A = LM.square_windows(lat), LM.sublattice_mask(lat)
    net.set_windows(b, W, A)
    ev, rep, _, u = q.lanczos(e, max_iter=300, want_eigvec=True)
    net.add_sample(b, u, 1.0, 0.0)
    theta = net.params().astype(np.float64)
    L, grad = net.loss_grad()
    reps, stab, _ = b.export()
    pats = OMLP.window_patterns(reps, 16, W)
    prior = OMLP.marshall_phase(reps, 16, A)
    L_or, g_or = OMLP.fidelity_loss(theta, pats, stab, 32, u, 1.0, 0.0, with_grad=True, round_like_gpu=True,
                                    prior=prior)
    assert abs(L - L_or) < 2e-3 * max(1.0, abs(L_or))
    assert grad @ g_or / (np.linalg.norm(grad) * np.linalg.norm(g_or)) > 0.999
    net.clear_samples()
    for (n1, n2, deltas) in ((4, 4, (0.8, 1.0, 1.2)), (4, 6, (0.8, 1.2))):
        for d in deltas:
            lat, b, e = sector(n1, n2, d)
            net.set_windows(b, LM.square_windows(lat), LM.sublattice_mask(lat))
            ev, rep, _, u = q.lanczos(e, max_iter=300, want_eigvec=True)
            net.add_sample(b, u, d, 0.0)
    L0, _ = net.loss_grad(want_grad=False)
    net.train(200, 3e-3)
    L1, _ = net.loss_grad(want_grad=False)
    print(f"square loss {L0:.4f} -> {L1:.4f}", flush=True)
    assert L1 < L0
    lat, b, e = sector(4, 6, 1.0)
    net.set_windows(b, LM.square_windows(lat), LM.sublattice_mask(lat))
    ev_r, rep_r, _, _ = q.lanczos(e, max_iter=300, tol=1e-13, check_every=1)
    ev_l, rep_l, fb, _ = net.lanczos(b, e, 1.0, 0.0, max_iter=300, tol=1e-13, check_every=1)
    print(f"4x6 iterations random {rep_r.iters}, learned {rep_l.iters}", flush=True)
    assert not fb and rep_l.converged
    assert abs(ev_l[0] - ev_r[0]) < 1e-10 * abs(ev_r[0])
    assert rep_l.iters < rep_r.iters
