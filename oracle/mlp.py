"""FP64 CPU oracle of the learned start vector (test infrastructure only).

The network (no reference counterpart): input = 8 spins of a window (+-1,
spin up = +1) plus Delta and J2/J1; two tanh hidden layers of 128; two
outputs (log-amplitude a, phase phi). For a basis state s of an N-site ring,
log psi(s) = sum_t a(w_t(s)) + i phi(w_t(s)) + i pi M(s), w_t(s) = the spins
of window t (a ring: sites t..t+7 mod N; a 2-D lattice: the windows it
registers), M(s) = number of up spins on sublattice A (Marshall's sign rule
for a bipartite lattice, a fixed prior; the network learns the phase relative
to it).
The k=0 symmetric-basis coefficient of representative r is
c_r = psi(r) * sqrt(G/m_r). Training loss per ground state u (real, unit):
L = 1 - |<u, c>|^2 / (<c, c> <u, u>).

Parameter vector layout (shared with the HIP library):
W1[128][10], b1[128], W2[128][128], b2[128], W3[2][128], b3[2].
The GPU rounds weights and the layer-2 inputs to bfloat16 (FP32 accumulate);
`bf16` below reproduces that rounding so tests can compare like with like.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT, WIN = 10, 128, 2, 8
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SIZES])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = ((u + 0x7FFF + lsb) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def pattern_inputs(delta, j2):
    """All 256 window patterns (bit i of p = spin at window offset i) with the couplings."""
    p = np.arange(256)
    spins = ((p[:, None] >> np.arange(WIN)) & 1) * 2.0 - 1.0
    return np.c_[spins, np.full(256, delta), np.full(256, j2)]


def forward(theta, X, round_like_gpu=True):
    P = unpack(theta)
    W2 = bf16(P["W2"]) if round_like_gpu else P["W2"]
    A1 = X @ P["W1"].T + P["b1"]
    H1 = np.tanh(A1)
    H1q = bf16(H1) if round_like_gpu else H1
    A2 = H1q @ W2.T + P["b2"]
    H2 = np.tanh(A2)
    out = H2 @ P["W3"].T + P["b3"]
    return out, (X, H1, H1q, H2, W2)


def backward(theta, cache, dout, round_like_gpu=True):
    P = unpack(theta)
    X, H1, H1q, H2, W2 = cache
    g = {}
    g["W3"] = dout.T @ H2
    g["b3"] = dout.sum(0)
    dA2 = (dout @ P["W3"]) * (1 - H2 ** 2)
    g["W2"] = dA2.T @ H1q
    g["b2"] = dA2.sum(0)
    dA1 = (dA2 @ W2) * (1 - H1 ** 2)
    g["W1"] = dA1.T @ X
    g["b1"] = dA1.sum(0)
    return pack(g)


def window_patterns(reps, n_sites, windows=None):
    """patterns[r, t] = pattern of window t of representative r: bit k = spin at
    site windows[t][k]; without windows, the ring windows t..t+7 (mod N)."""
    if windows is None:
        windows = [[(t + k) % n_sites for k in range(WIN)] for t in range(n_sites)]
    windows = np.asarray(windows, np.int64)
    reps = np.asarray(reps, np.uint64)
    out = np.zeros((len(reps), len(windows)), np.int64)
    for k in range(WIN):
        out |= ((reps[:, None] >> windows[None, :, k].astype(np.uint64)) & np.uint64(1)).astype(np.int64) << k
    return out


def marshall_phase(reps, n_sites, sublattice_a=None):
    """pi * (up spins on sublattice A) of each representative; A defaults to the
    even sites of a ring."""
    if sublattice_a is None:
        sublattice_a = sum(1 << j for j in range(0, n_sites, 2))
    return np.array([np.pi * bin(int(r) & sublattice_a).count("1") for r in np.asarray(reps, np.uint64)],
                    np.float64)


def log_psi(theta, pats, prior, delta, j2, round_like_gpu=True):
    """(alpha_r, beta_r): log-amplitude and phase of each representative."""
    out, cache = forward(theta, pattern_inputs(delta, j2), round_like_gpu)
    alpha = out[:, 0][pats].sum(1)
    beta = out[:, 1][pats].sum(1) + (0.0 if prior is None else prior)
    return alpha, beta, cache


def start_vector(theta, pats, prior, stab, G, delta, j2):
    """Real start vector |c_r| cos(beta_r - beta_ref), beta_ref = phase of the
    first representative of largest amplitude (before the noise admixture)."""
    alpha, beta, _ = log_psi(theta, pats, prior, delta, j2)
    i = int(np.argmax(alpha))
    return np.exp(alpha - alpha[i]) * np.sqrt(G / np.asarray(stab, np.float64)) * np.cos(beta - beta[i])


def start_noise(n, seed):
    """xi_r = 2 u_r - 1 with u_r element r of the SplitMix64 stream of seed."""
    m = (1 << 64) - 1
    out = np.empty(n, np.float64)
    for c in range(n):
        z = (seed + (c + 1) * 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        z ^= z >> 31
        out[c] = 2.0 * (((z >> 11) + 1) * 2.0 ** -53) - 1.0
    return out


def fidelity_loss(theta, pats, stab, G, u, delta, j2, with_grad=False, round_like_gpu=True, prior=None):
    """Loss 1 - F for one ground state; optional gradient wrt theta (FP64).
    prior: the fixed phase per representative (marshall_phase), or None."""
    alpha, beta, cache = log_psi(theta, pats, prior, delta, j2, round_like_gpu)
    amax = alpha.max()
    c = np.exp(alpha - amax + 1j * beta) * np.sqrt(G / np.asarray(stab, np.float64))
    O = np.sum(u * c)
    Nc = np.sum(np.abs(c) ** 2)
    Nu = np.sum(u * u)
    F = abs(O) ** 2 / (Nc * Nu)
    if not with_grad:
        return 1.0 - F
    # dF/d(alpha_r), dF/d(beta_r)
    w = np.conj(O) * u * c
    g_alpha = (2 * w.real * Nc - abs(O) ** 2 * 2 * np.abs(c) ** 2) / (Nc ** 2 * Nu)
    g_beta = (-2 * w.imag) / (Nc * Nu)
    dout = np.zeros((256, 2))
    np.add.at(dout[:, 0], pats.reshape(-1), np.repeat(-g_alpha, pats.shape[1]))
    np.add.at(dout[:, 1], pats.reshape(-1), np.repeat(-g_beta, pats.shape[1]))
    return 1.0 - F, backward(theta, cache, dout, round_like_gpu)


def init_params(seed):
    """Deterministic FP32 initialisation (uniform, fan-in scaled). The library
    generates the same numbers with the same SplitMix64 stream."""
    vals = []
    c = 0
    fan = {"W1": IN, "b1": IN, "W2": H, "b2": H, "W3": H, "b3": H}
    m = (1 << 64) - 1
    for name, shape in SIZES:
        n = int(np.prod(shape))
        lim = 1.0 / np.sqrt(fan[name])
        for _ in range(n):
            z = (seed + (c + 1) * 0x9E3779B97F4A7C15) & m
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
            z ^= z >> 31
            u01 = ((z >> 11) + 1) * 2.0 ** -53
            vals.append(np.float32((2 * u01 - 1) * lim))
            c += 1
    return np.array(vals, np.float64)
