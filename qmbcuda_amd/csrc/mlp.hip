// Learned Lanczos start vector (no reference counterpart).
//
// Network: input = 8 spins of a window (+-1) + Delta + J2/J1 (10 values);
// tanh(128) -> tanh(128) -> (log-amplitude a, phase phi). For an N-site ring
// state s: log psi(s) = sum_t a(w_t) + i phi(w_t), w_t = spins t..t+7 (mod N),
// translation invariant by construction. The k=0 coefficient of representative
// r is c_r = psi(r) sqrt(G/m_r).
//
// Because a window takes only 256 values, the network is evaluated once per
// (pattern, coupling) row; states then cost N table lookups. Training
// back-propagates per-pattern gradient histograms. The 128x128 layer - the
// bulk of the arithmetic - runs on gfx950 MFMA (v_mfma_f32_32x32x16_bf16:
// bf16 weights and activations, FP32 accumulation) in the forward pass, the
// activation backward pass (dH1 = dA2 W2) and the weight gradient
// (dW2 = dA2^T H1); the 10-wide input layer and the 2-wide head use FP32 VALU.
// Weight-gradient partial sums, the pattern histograms (fixed-point int64) and
// every reduction are deterministic.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>

#include "internal.hpp"
#include "reduce.hpp"

namespace qmb {

constexpr int kIn = 10, kH = 128, kOut = 2, kWin = 8;
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kOut * kH, kParams = ob3 + kOut;
constexpr double kFix = 68719476736.0;  // 2^36 fixed-point scale of the pattern histograms
constexpr double kMarshallPi = 3.14159265358979323846;
constexpr double kStartNoise = 0.1;  // weight of the unit seeded noise in the Lanczos start

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline uint16_t to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16_bits(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline bf16x8 load8(const uint16_t *p) {
    bf16x8 v;
    __builtin_memcpy(&v, p, 16);
    return v;
}

// C/D layout of the 32x32 MFMA tile held by lane l: element v sits at
// row 8*(v/4) + 4*(l/32) + (v%4), column l%32. A fragment: row l%32, k = 8*(l/32)..+7;
// B fragment: column l%32, k = 8*(l/32)..+7 (16 k per instruction).
__device__ inline int mfma_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// ---- forward: one wavefront per 32 rows --------------------------------------
// Row r = c*256 + p: pattern p, coupling c. Writes out[r][2]; with store, also
// H1 (FP32 and bf16) and H2 (FP32) for the backward pass.
constexpr int kTileLd = kH + 8;  // bf16 row stride of the LDS activation tile

__global__ __launch_bounds__(64) void mlp_forward(const float *__restrict__ theta, const uint16_t *__restrict__ w2bf,
                                                  const double *__restrict__ coup, int n_rows,
                                                  float *__restrict__ out, float *__restrict__ H1,
                                                  uint16_t *__restrict__ H1q, float *__restrict__ H2, int store) {
    __shared__ float x[32][kIn];
    __shared__ uint16_t a_tile[32 * kTileLd];
    __shared__ float h2[32][kH + 1];
    const int lane = threadIdx.x;
    const int r0 = blockIdx.x * 32;
    if (lane < 32) {
        const int r = r0 + lane;
        const int p = r & 255;
        const int c = r >> 8;
        for (int i = 0; i < kWin; ++i) x[lane][i] = ((p >> i) & 1) ? 1.f : -1.f;
        x[lane][8] = r < n_rows ? (float)coup[2 * c] : 0.f;
        x[lane][9] = r < n_rows ? (float)coup[2 * c + 1] : 0.f;
    }
    __syncthreads();
    // layer 1 (VALU): lane owns hidden units 2*lane, 2*lane+1
    for (int u = 0; u < 2; ++u) {
        const int k = 2 * lane + u;
        float w[kIn];
        for (int i = 0; i < kIn; ++i) w[i] = theta[oW1 + k * kIn + i];
        const float b = theta[ob1 + k];
        for (int row = 0; row < 32; ++row) {
            float s = b;
            for (int i = 0; i < kIn; ++i) s += w[i] * x[row][i];
            const float h = tanhf(s);
            const uint16_t q = to_bf16_bits(h);
            a_tile[row * kTileLd + k] = q;
            const int r = r0 + row;
            if (store && r < n_rows) {
                H1[(size_t)r * kH + k] = h;
                H1q[(size_t)r * kH + k] = q;
            }
        }
    }
    __syncthreads();
    // layer 2 (MFMA): H2 = tanh(H1q W2^T + b2)
    for (int nt = 0; nt < 4; ++nt) {
        f32x16 acc = {};
        for (int kb = 0; kb < kH / 16; ++kb) {
            const bf16x8 a = load8(a_tile + (lane & 31) * kTileLd + kb * 16 + 8 * (lane >> 5));
            const bf16x8 b = load8(w2bf + (size_t)(nt * 32 + (lane & 31)) * kH + kb * 16 + 8 * (lane >> 5));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = nt * 32 + (lane & 31);
        const float b2 = theta[ob2 + col];
        for (int v = 0; v < 16; ++v) h2[mfma_row(v, lane)][col] = tanhf(acc[v] + b2);
    }
    __syncthreads();
    // head (VALU): lane -> (row = lane%32, output = lane/32)
    {
        const int row = lane & 31, o = lane >> 5;
        float s = theta[ob3 + o];
        for (int k = 0; k < kH; ++k) s += theta[oW3 + o * kH + k] * h2[row][k];
        const int r = r0 + row;
        if (r < n_rows) out[(size_t)r * 2 + o] = s;
    }
    if (store) {
        for (int idx = lane; idx < 32 * kH; idx += 64) {
            const int row = idx / kH, k = idx % kH;
            const int r = r0 + row;
            if (r < n_rows) H2[(size_t)r * kH + k] = h2[row][k];
        }
    }
}

// ---- backward, activations: dA2 = (dout W3) (1 - H2^2); dA1 = (dA2 W2) (1 - H1^2)
// dA1 uses MFMA with B = W2 read column-wise through its transpose w2t[j][n].
__global__ __launch_bounds__(64) void mlp_backward_act(const float *__restrict__ theta,
                                                       const uint16_t *__restrict__ w2t, int n_rows,
                                                       const float *__restrict__ dout, const float *__restrict__ H1,
                                                       const float *__restrict__ H2, float *__restrict__ dA2,
                                                       uint16_t *__restrict__ dA2q, float *__restrict__ dA1) {
    __shared__ uint16_t a_tile[32 * kTileLd];
    const int lane = threadIdx.x;
    const int r0 = blockIdx.x * 32;
    for (int idx = lane; idx < 32 * kH; idx += 64) {
        const int row = idx / kH, k = idx % kH;
        const int r = r0 + row;
        float g = 0.f;
        if (r < n_rows) {
            const float h = H2[(size_t)r * kH + k];
            g = (dout[(size_t)r * 2] * theta[oW3 + k] + dout[(size_t)r * 2 + 1] * theta[oW3 + kH + k]) * (1.f - h * h);
            dA2[(size_t)r * kH + k] = g;
        }
        const uint16_t q = to_bf16_bits(g);
        a_tile[row * kTileLd + k] = q;
        if (r < n_rows) dA2q[(size_t)r * kH + k] = q;
    }
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt) {
        f32x16 acc = {};
        for (int kb = 0; kb < kH / 16; ++kb) {
            const bf16x8 a = load8(a_tile + (lane & 31) * kTileLd + kb * 16 + 8 * (lane >> 5));
            const bf16x8 b = load8(w2t + (size_t)(nt * 32 + (lane & 31)) * kH + kb * 16 + 8 * (lane >> 5));
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = nt * 32 + (lane & 31);
        for (int v = 0; v < 16; ++v) {
            const int r = r0 + mfma_row(v, lane);
            if (r < n_rows) {
                const float h = H1[(size_t)r * kH + col];
                dA1[(size_t)r * kH + col] = acc[v] * (1.f - h * h);
            }
        }
    }
}

// ---- backward, dW2 = dA2^T H1q over a group of rows (MFMA, K = rows) ---------
// grid (groups, 16): blockIdx.y = output tile (mt, nt); one partial tile per group.
constexpr int kRowsPerGroup = 1024;

__global__ __launch_bounds__(64) void mlp_wgrad2(int n_rows, const uint16_t *__restrict__ dA2q,
                                                 const uint16_t *__restrict__ H1q, float *__restrict__ part) {
    __shared__ uint16_t at[32][40];  // dA2^T tile: [m][row]
    __shared__ uint16_t bt[32][40];  // H1q^T tile: [n][row]
    const int lane = threadIdx.x;
    const int mt = blockIdx.y >> 2, nt = blockIdx.y & 3;
    const int g0 = blockIdx.x * kRowsPerGroup;
    f32x16 acc = {};
    for (int t0 = g0; t0 < g0 + kRowsPerGroup && t0 < n_rows; t0 += 32) {
        for (int idx = lane; idx < 32 * 32; idx += 64) {
            const int row = idx / 32, m = idx % 32;
            const int r = t0 + row;
            at[m][row] = r < n_rows ? dA2q[(size_t)r * kH + mt * 32 + m] : (uint16_t)0;
            bt[m][row] = r < n_rows ? H1q[(size_t)r * kH + nt * 32 + m] : (uint16_t)0;
        }
        __syncthreads();
        for (int kb = 0; kb < 2; ++kb) {
            const bf16x8 a = load8(&at[lane & 31][kb * 16 + 8 * (lane >> 5)]);
            const bf16x8 b = load8(&bt[lane & 31][kb * 16 + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    float *dst = part + ((size_t)blockIdx.x * kH * kH);
    for (int v = 0; v < 16; ++v) {
        const int m = mt * 32 + mfma_row(v, lane), n = nt * 32 + (lane & 31);
        dst[(size_t)m * kH + n] = acc[v];
    }
}

// ---- backward, small parameters (VALU) over a group of rows ----------------
// per group partial of W1, b1, b2, W3, b3 (W2 comes from mlp_wgrad2). Groups of
// 256 rows (one coupling's patterns): four times as many blocks as the MFMA
// weight gradient's groups, so the row loops are short.
constexpr int kRowsPerSmallGroup = 256;

__global__ __launch_bounds__(kH) void mlp_wgrad_small(const double *__restrict__ coup, int n_rows,
                                                      const float *__restrict__ dout, const float *__restrict__ H2,
                                                      const float *__restrict__ dA2, const float *__restrict__ dA1,
                                                      float *__restrict__ part) {
    const int k = threadIdx.x;  // hidden unit
    const int g0 = blockIdx.x * kRowsPerSmallGroup;
    float gw1[kIn] = {}, gb1 = 0.f, gb2 = 0.f, gw3a = 0.f, gw3b = 0.f, gb3a = 0.f, gb3b = 0.f;
    for (int r = g0; r < g0 + kRowsPerSmallGroup && r < n_rows; ++r) {
        const int p = r & 255, c = r >> 8;
        const float d1 = dA1[(size_t)r * kH + k];
        for (int i = 0; i < kWin; ++i) gw1[i] += d1 * (((p >> i) & 1) ? 1.f : -1.f);
        gw1[8] += d1 * (float)coup[2 * c];
        gw1[9] += d1 * (float)coup[2 * c + 1];
        gb1 += d1;
        gb2 += dA2[(size_t)r * kH + k];
        const float h2 = H2[(size_t)r * kH + k];
        gw3a += dout[(size_t)r * 2] * h2;
        gw3b += dout[(size_t)r * 2 + 1] * h2;
        if (k == 0) {
            gb3a += dout[(size_t)r * 2];
            gb3b += dout[(size_t)r * 2 + 1];
        }
    }
    float *dst = part + (size_t)blockIdx.x * kParams;
    for (int i = 0; i < kIn; ++i) dst[oW1 + k * kIn + i] = gw1[i];
    dst[ob1 + k] = gb1;
    dst[ob2 + k] = gb2;
    dst[oW3 + k] = gw3a;
    dst[oW3 + kH + k] = gw3b;
    if (k == 0) {
        dst[ob3] = gb3a;
        dst[ob3 + 1] = gb3b;
    }
}

// grad = scale * sum over groups of the partials (W2 from part2 over n_groups2
// MFMA groups, the rest from part over n_groups small groups).
__global__ __launch_bounds__(256) void mlp_reduce_grad(int n_groups, int n_groups2, const float *__restrict__ part,
                                                       const float *__restrict__ part2, float scale,
                                                       float *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float s = 0.f;
    if (i >= oW2 && i < ob2) {
        for (int g = 0; g < n_groups2; ++g) s += part2[(size_t)g * kH * kH + (i - oW2)];
    } else {
        for (int g = 0; g < n_groups; ++g) s += part[(size_t)g * kParams + i];
    }
    grad[i] = s * scale;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refresh the bf16 W2 and its transpose.
__global__ __launch_bounds__(256) void mlp_adam(float *__restrict__ theta, const float *__restrict__ grad,
                                                float *__restrict__ m1, float *__restrict__ m2, float lr, float b1,
                                                float b2, float bc1, float bc2, uint16_t *__restrict__ w2bf,
                                                uint16_t *__restrict__ w2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * g;
    const float v = b2 * m2[i] + (1.f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    const float t = theta[i] - lr * (a / bc1) / (sqrtf(v / bc2) + 1e-8f);
    theta[i] = t;
    if (i >= oW2 && i < ob2) {
        const int n = (i - oW2) / kH, k = (i - oW2) % kH;
        const uint16_t q = to_bf16_bits(t);
        w2bf[n * kH + k] = q;
        w2t[k * kH + n] = q;
    }
}

__global__ __launch_bounds__(256) void mlp_refresh_w2(const float *__restrict__ theta, uint16_t *__restrict__ w2bf,
                                                      uint16_t *__restrict__ w2t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kH * kH) return;
    const int n = i / kH, k = i % kH;
    const uint16_t q = to_bf16_bits(theta[oW2 + i]);
    w2bf[n * kH + k] = q;
    w2t[k * kH + n] = q;
}

// ---- per-state kernels (one sample = one ground state) ---------------------
// A window is 8 site indices; its pattern has bit k = spin at site k of the
// window. Rings use windows q..q+7 (mod N); 2-D lattices register their own.
constexpr int kMaxWin = 128;

__device__ inline int window_pattern(uint64_t s, const uint8_t *w) {
    int p = 0;
#pragma unroll
    for (int k = 0; k < kWin; ++k) p |= (int)((s >> w[k]) & 1ull) << k;
    return p;
}

__device__ inline void stage_windows(uint8_t *lds, const uint8_t *win, int n_win) {
    for (int i = threadIdx.x; i < n_win * kWin; i += blockDim.x) lds[i] = win[i];
}

// alpha_r, beta_r from the pattern table of this coupling; block max of alpha.
// beta_r also carries the Marshall sign prior pi * (up spins on sublattice A):
// the network learns the phase relative to the bipartite sign rule, whose
// state (and not the uniform one, which is the fully polarised multiplet and
// orthogonal to a singlet) is the natural starting point of training.
__global__ __launch_bounds__(256) void mlp_state_logpsi(int64_t n, int n_win, const uint8_t *__restrict__ win,
                                                        uint64_t sub_a, const uint64_t *__restrict__ reps,
                                                        const float *__restrict__ tab, double *__restrict__ alpha,
                                                        double *__restrict__ beta, double *__restrict__ part) {
    __shared__ float t[512];
    __shared__ double lds[kWaves];
    __shared__ uint8_t wl[kMaxWin * kWin];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) t[i] = tab[i];
    stage_windows(wl, win, n_win);
    __syncthreads();
    double mx = -1e300;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t s = reps[r];
        double a = 0.0, b = 0.0;
        for (int q = 0; q < n_win; ++q) {
            const int p = window_pattern(s, wl + q * kWin);
            a += (double)t[2 * p];
            b += (double)t[2 * p + 1];
        }
        alpha[r] = a;
        beta[r] = b + kMarshallPi * (double)__popcll(s & sub_a);
        mx = fmax(mx, a);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int i = 1; i < kWaves; ++i) m = fmax(m, lds[i]);
        part[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(256) void mlp_finish_max(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double lds[kWaves];
    double mx = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, part[i]);
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int i = 1; i < kWaves; ++i) m = fmax(m, lds[i]);
        *out = m;
    }
}

// Real start vector x_r = |c_r| cos(beta_r - beta_ref) with beta_ref the phase
// of the first representative of largest amplitude (alpha == amax). Block
// partials: part[block] = sum of x_r^2, part[gridDim.x + block] = sum of xi_r^2
// for the seeded noise xi_r = 2 u_r - 1 (u_r = element r of the SplitMix64
// stream of `seed`) that the Lanczos start mixes in.
__global__ __launch_bounds__(256) void mlp_state_start(int64_t n, int G, const uint8_t *__restrict__ stab,
                                                       const double *__restrict__ alpha,
                                                       const double *__restrict__ beta, const double *__restrict__ amax,
                                                       const double *__restrict__ beta_ref, uint64_t seed,
                                                       double *__restrict__ x, double *__restrict__ part) {
    __shared__ double lds[kWaves];
    const double am = *amax, br = *beta_ref;
    double acc = 0.0, accn = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double mag = exp(alpha[r] - am) * sqrt((double)G / (double)stab[r]);
        const double v = mag * cos(beta[r] - br);
        x[r] = v;
        acc += v * v;
        const double xi = 2.0 * splitmix_uniform(seed, (uint64_t)r) - 1.0;
        accn += xi * xi;
    }
    const double t = block_sum(acc, lds);
    const double tn = block_sum(accn, lds);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = t;
        part[gridDim.x + blockIdx.x] = tn;
    }
}

// x_r <- x_r / ||x|| + eta * xi_r / ||xi|| (norms squared in sums[0], sums[1]).
__global__ __launch_bounds__(256) void mlp_mix_noise(int64_t n, uint64_t seed, double eta,
                                                     const double *__restrict__ sums, double *__restrict__ x) {
    const double ix = 1.0 / sqrt(sums[0]), in = eta / sqrt(sums[1]);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const double xi = 2.0 * splitmix_uniform(seed, (uint64_t)r) - 1.0;
        x[r] = x[r] * ix + xi * in;
    }
}

// Smallest index whose alpha equals amax: per-block minima (part[block] =
// index as a double, exact below 2^53), then one block picks the minimum and
// stores beta at that index.
__device__ inline long long block_min_ll(long long v, long long *lds) {
    for (int off = 32; off > 0; off >>= 1) {
        const long long o = __shfl_down(v, off, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    long long b = lds[0];
    for (int i = 1; i < kWaves; ++i) b = lds[i] < b ? lds[i] : b;
    __syncthreads();
    return b;
}

__global__ __launch_bounds__(256) void mlp_argmax_part(int64_t n, const double *__restrict__ alpha,
                                                       const double *__restrict__ amax, double *__restrict__ part) {
    __shared__ long long lds[kWaves];
    const double am = *amax;
    long long best = LLONG_MAX;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
        if (alpha[r] == am && r < best) best = r;
    const long long b = block_min_ll(best, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = b == LLONG_MAX ? -1.0 : (double)b;
}

__global__ __launch_bounds__(256) void mlp_beta_ref(int nb, const double *__restrict__ part,
                                                    const double *__restrict__ beta, double *__restrict__ beta_ref) {
    __shared__ long long lds[kWaves];
    long long best = LLONG_MAX;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const double v = part[i];
        if (v >= 0.0 && (long long)v < best) best = (long long)v;
    }
    const long long b = block_min_ll(best, lds);
    if (threadIdx.x == 0) *beta_ref = b == LLONG_MAX ? 0.0 : beta[b];
}

// out[0], out[1] = fixed-order sums of part[0..n) and part[n..2n).
__global__ __launch_bounds__(256) void mlp_finish_sum2(const double *__restrict__ part, int n,
                                                       double *__restrict__ out) {
    __shared__ double lds[kWaves];
    const double t0 = sum_partials(part, n, lds);
    const double t1 = sum_partials(part + n, n, lds);
    if (threadIdx.x == 0) {
        out[0] = t0;
        out[1] = t1;
    }
}

// ---- the per-sample kernels batched over every sample of a training step ----
// One launch per phase covers all samples: block b works on sample task[b],
// with the row-to-thread assignment and the per-sample block count of the
// single-sample kernel above, so every sample's sums are formed in a fixed order.
struct BSample {
    int64_t n, off;  // rows; offset of the sample's rows in the concatenated alpha/beta
    const uint64_t *reps;
    const uint8_t *stab;
    const double *u;
    const uint8_t *win;
    uint64_t sub_a;
    int n_win, G, blk0, nblk;  // blocks [blk0, blk0 + nblk) of every batched launch
};

__global__ __launch_bounds__(256) void mlp_batch_logpsi(const BSample *__restrict__ S, const int *__restrict__ task,
                                                        const float *__restrict__ tabs, double *__restrict__ alpha,
                                                        double *__restrict__ beta, double *__restrict__ part) {
    __shared__ float t[512];
    __shared__ double lds[kWaves];
    __shared__ uint8_t wl[kMaxWin * kWin];
    const int si = task[blockIdx.x];
    const BSample s = S[si];
    const float *tab = tabs + (size_t)si * 512;
    for (int i = threadIdx.x; i < 512; i += blockDim.x) t[i] = tab[i];
    stage_windows(wl, s.win, s.n_win);
    __syncthreads();
    const int lb = blockIdx.x - s.blk0;
    double mx = -1e300;
    for (int64_t r = (int64_t)lb * blockDim.x + threadIdx.x; r < s.n; r += (int64_t)s.nblk * blockDim.x) {
        const uint64_t st = s.reps[r];
        double a = 0.0, b = 0.0;
        for (int q = 0; q < s.n_win; ++q) {
            const int p = window_pattern(st, wl + q * kWin);
            a += (double)t[2 * p];
            b += (double)t[2 * p + 1];
        }
        alpha[s.off + r] = a;
        beta[s.off + r] = b + kMarshallPi * (double)__popcll(st & s.sub_a);
        mx = fmax(mx, a);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int i = 1; i < kWaves; ++i) m = fmax(m, lds[i]);
        part[blockIdx.x] = m;
    }
}

// One block per sample: stats[8 si + 5] = max of the sample's block maxima.
__global__ __launch_bounds__(256) void mlp_batch_finish_max(const BSample *__restrict__ S,
                                                            const double *__restrict__ part,
                                                            double *__restrict__ stats) {
    __shared__ double lds[kWaves];
    const BSample s = S[blockIdx.x];
    double mx = -1e300;
    for (int i = threadIdx.x; i < s.nblk; i += blockDim.x) mx = fmax(mx, part[s.blk0 + i]);
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lds[0];
        for (int i = 1; i < kWaves; ++i) m = fmax(m, lds[i]);
        stats[(size_t)blockIdx.x * 8 + 5] = m;
    }
}

// Partials of Re/Im <u, c>, <c, c>, <u, u>: part[q * total + block].
__global__ __launch_bounds__(256) void mlp_batch_overlap(const BSample *__restrict__ S, const int *__restrict__ task,
                                                         const double *__restrict__ alpha,
                                                         const double *__restrict__ beta,
                                                         const double *__restrict__ stats, double *__restrict__ part,
                                                         int total) {
    __shared__ double lds[kWaves];
    const int si = task[blockIdx.x];
    const BSample s = S[si];
    const double am = stats[(size_t)si * 8 + 5];
    const int lb = blockIdx.x - s.blk0;
    double o_re = 0, o_im = 0, nc = 0, nu = 0;
    for (int64_t r = (int64_t)lb * blockDim.x + threadIdx.x; r < s.n; r += (int64_t)s.nblk * blockDim.x) {
        const double mag = exp(alpha[s.off + r] - am) * sqrt((double)s.G / (double)s.stab[r]);
        const double br = beta[s.off + r];
        const double cr = mag * cos(br), ci = mag * sin(br);
        const double ur = s.u[r];
        o_re += ur * cr;
        o_im += ur * ci;
        nc += mag * mag;
        nu += ur * ur;
    }
    const double v[4] = {o_re, o_im, nc, nu};
    for (int q = 0; q < 4; ++q) {
        const double t = block_sum(v[q], lds);
        if (threadIdx.x == 0) part[(size_t)q * total + blockIdx.x] = t;
    }
}

// One block per sample: stats[8 si + 0..3] = Re O, Im O, Nc, Nu; stats[8 si + 4] = loss.
__global__ __launch_bounds__(256) void mlp_batch_finish_overlap(const BSample *__restrict__ S,
                                                                const double *__restrict__ part, int total,
                                                                double *__restrict__ stats) {
    __shared__ double lds[kWaves];
    const BSample s = S[blockIdx.x];
    double v[4];
    for (int q = 0; q < 4; ++q) v[q] = sum_partials(part + (size_t)q * total + s.blk0, s.nblk, lds);
    if (threadIdx.x == 0) {
        double *st = stats + (size_t)blockIdx.x * 8;
        for (int q = 0; q < 4; ++q) st[q] = v[q];
        const double o2 = v[0] * v[0] + v[1] * v[1];
        st[4] = 1.0 - o2 / (v[2] * v[3]);
    }
}

__global__ __launch_bounds__(256) void mlp_zero_u64(unsigned long long *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Per-pattern gradient histograms (fixed point) of every sample: each block adds
// its LDS histogram into its sample's hist[512] (integer sums, any order).
__global__ __launch_bounds__(256) void mlp_batch_grad(const BSample *__restrict__ S, const int *__restrict__ task,
                                                      const double *__restrict__ alpha,
                                                      const double *__restrict__ beta,
                                                      const double *__restrict__ stats,
                                                      unsigned long long *__restrict__ hist) {
    __shared__ unsigned long long h[512];
    __shared__ uint8_t wl[kMaxWin * kWin];
    const int si = task[blockIdx.x];
    const BSample s = S[si];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) h[i] = 0ull;
    stage_windows(wl, s.win, s.n_win);
    __syncthreads();
    const double *st = stats + (size_t)si * 8;
    const double am = st[5];
    const double ore = st[0], oim = st[1], nc = st[2], nu = st[3];
    const double o2 = ore * ore + oim * oim;
    const int lb = blockIdx.x - s.blk0;
    const int lane = threadIdx.x & 63;
    // wave-uniform loop (lane l of a wave handles row r0 + l, the same rows as a
    // thread-strided loop): consecutive representatives share their high sites, so
    // a window there often has one pattern across the wave; its 64 additions are
    // then summed in registers and added once instead of contending for one bin
    for (int64_t r0 = (int64_t)lb * blockDim.x + (threadIdx.x & ~63); r0 < s.n; r0 += (int64_t)s.nblk * blockDim.x) {
        const int64_t r = r0 + lane;
        const bool act = r < s.n;
        long long qa = 0, qb = 0;
        uint64_t stt = 0;
        if (act) {
            const double mag = exp(alpha[s.off + r] - am) * sqrt((double)s.G / (double)s.stab[r]);
            const double br = beta[s.off + r];
            const double cr = mag * cos(br), ci = mag * sin(br);
            const double wr = s.u[r] * (ore * cr + oim * ci);
            const double wi = s.u[r] * (ore * ci - oim * cr);
            const double ga = -(2.0 * wr * nc - o2 * 2.0 * mag * mag) / (nc * nc * nu);
            const double gb = -(-2.0 * wi) / (nc * nu);
            qa = (long long)llrint(ga * kFix);
            qb = (long long)llrint(gb * kFix);
            stt = s.reps[r];
        }
        const uint64_t st0 = __shfl(stt, 0, 64);  // lane 0's row exists (r0 < n)
        if (!act) stt = st0;                     // rows past the end add 0 to lane 0's bins
        for (int q = 0; q < s.n_win; ++q) {
            const int p = window_pattern(stt, wl + q * kWin);
            const int p0 = __shfl(p, 0, 64);
            if (__all(p == p0)) {
                long long sa = qa, sb = qb;
                for (int off = 32; off > 0; off >>= 1) {
                    sa += __shfl_xor(sa, off, 64);
                    sb += __shfl_xor(sb, off, 64);
                }
                if (lane == 0) {
                    atomicAdd(&h[2 * p0], (unsigned long long)sa);
                    atomicAdd(&h[2 * p0 + 1], (unsigned long long)sb);
                }
            } else {
                atomicAdd(&h[2 * p], (unsigned long long)qa);
                atomicAdd(&h[2 * p + 1], (unsigned long long)qb);
            }
        }
    }
    __syncthreads();
    unsigned long long *hs = hist + (size_t)si * 512;
    for (int i = threadIdx.x; i < 512; i += blockDim.x)
        if (h[i]) atomicAdd(&hs[i], h[i]);
}

// dout rows of every sample (512 per sample) from its histogram; loss = mean of
// the samples' losses summed in sample order.
__global__ __launch_bounds__(256) void mlp_batch_finish(const unsigned long long *__restrict__ hist, int ns,
                                                        float *__restrict__ dout, float scale,
                                                        const double *__restrict__ stats, double *__restrict__ loss) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)ns * 512;
         i += (int64_t)gridDim.x * blockDim.x)
        dout[i] = (float)((double)(long long)hist[i] / kFix * scale);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double mean = 0.0;
        for (int k = 0; k < ns; ++k) mean += stats[(size_t)k * 8 + 4];
        *loss = mean / ns;
    }
}

// Loss only: the mean of the samples' losses.
__global__ __launch_bounds__(64) void mlp_batch_loss(const double *__restrict__ stats, int ns,
                                                     double *__restrict__ loss) {
    if (threadIdx.x != 0) return;
    double mean = 0.0;
    for (int k = 0; k < ns; ++k) mean += stats[(size_t)k * 8 + 4];
    *loss = mean / ns;
}

}  // namespace qmb

using namespace qmb;

// Window geometry of one basis: n_win windows of 8 sites and the sublattice A
// mask of the Marshall prior.
struct MlpGeom {
    Ref<const qmb_basis> basis;
    int n_win = 0;
    DevBuf win;  // uint8 [n_win * 8]
    uint64_t sub_a = 0;
};

struct MlpSample {
    Ref<const qmb_basis> basis;
    int geom = 0;  // index into qmb_mlp::geoms
    DevBuf u;      // target ground state
    double delta, j2;
};

// Device tables and work arrays of the batched per-sample kernels, rebuilt when
// the sample list changes.
struct MlpBatch {
    bool valid = false;
    int ns = 0, total_blocks = 0;
    int64_t total_rows = 0;
    DevBuf samples, task, alpha, beta, part, stats, hist, loss;
};

struct qmb_mlp : Counted {
    Ref<qmb_ctx> ctx;
    DevBuf theta, m1, m2, grad, w2bf, w2t;
    int64_t step = 0;
    std::vector<MlpGeom> geoms;
    std::vector<MlpSample> samples;
    MlpBatch batch;
    // row buffers (sized for the current row count)
    int rows_cap = 0;
    DevBuf coup, out, H1, H1q, H2, dout, dA2, dA2q, dA1, part, part2;
    void ensure_rows(int rows);
};

void qmb_mlp::ensure_rows(int rows) {
    if (rows <= rows_cap) return;
    const int cap = ((rows + 255) / 256) * 256;
    coup.alloc((size_t)(cap / 256) * 2 * sizeof(double));
    out.alloc((size_t)cap * 2 * sizeof(float));
    H1.alloc((size_t)cap * kH * sizeof(float));
    H1q.alloc((size_t)cap * kH * sizeof(uint16_t));
    H2.alloc((size_t)cap * kH * sizeof(float));
    dout.alloc((size_t)cap * 2 * sizeof(float));
    dA2.alloc((size_t)cap * kH * sizeof(float));
    dA2q.alloc((size_t)cap * kH * sizeof(uint16_t));
    dA1.alloc((size_t)cap * kH * sizeof(float));
    const int groups = (cap + kRowsPerGroup - 1) / kRowsPerGroup;
    const int groups_s = (cap + kRowsPerSmallGroup - 1) / kRowsPerSmallGroup;
    part.alloc((size_t)groups_s * kParams * sizeof(float));
    part2.alloc((size_t)groups * kH * kH * sizeof(float));
    rows_cap = cap;
}

namespace {

// The network's bases: compact layout, every character 1 (k = 0, and
// flip-even when the group has the spin flip).
void check_mlp_basis(const qmb_basis *b) {
    QMB_REQUIRE(b->layout == QMB_LAYOUT_COMPACT, "mlp: the basis must use the compact layout");
    QMB_REQUIRE(b->all_rows, "mlp: the basis must be the fully symmetric sector (k = 0, flip-even)");
    QMB_REQUIRE(b->n_sites >= kWin && b->n_sites <= 64, "mlp: lattice size must be in [8, 64]");
}

// Index of the geometry of basis b: registered windows, or for a plain ring
// (translation group G = N) the N windows q..q+7 (mod N) with sublattice A =
// the even sites.
int geom_for(qmb_mlp *m, const qmb_basis *b) {
    for (size_t i = 0; i < m->geoms.size(); ++i)
        if (m->geoms[i].basis == b) return (int)i;
    check_mlp_basis(b);
    const qmb_group *g = b->group;
    QMB_REQUIRE(g->torus && g->block == g->n_sites && g->G == g->n_sites && b->n_sites <= kMaxWin,
                "mlp: no windows registered for this basis and it is not a plain ring (G = N)");
    const int N = b->n_sites;
    std::vector<uint8_t> w((size_t)N * kWin);
    for (int q = 0; q < N; ++q)
        for (int k = 0; k < kWin; ++k) w[(size_t)q * kWin + k] = (uint8_t)((q + k) % N);
    MlpGeom gm;
    gm.basis = b;
    gm.n_win = N;
    gm.win.alloc(w.size());
    QMB_HIP(hipMemcpy(gm.win.p, w.data(), w.size(), hipMemcpyHostToDevice));
    for (int q = 0; q < N; q += 2) gm.sub_a |= 1ULL << q;
    m->geoms.push_back(std::move(gm));
    return (int)m->geoms.size() - 1;
}

// Forward pass over the given couplings (rows = 256 per coupling).
void forward_rows(qmb_mlp *m, const std::vector<double> &coup, bool store) {
    qmb_ctx *ctx = m->ctx;
    const int rows = (int)(coup.size() / 2) * 256;
    m->ensure_rows(rows);
    QMB_HIP(hipMemcpyAsync(m->coup.p, coup.data(), coup.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mlp_forward, dim3((rows + 31) / 32), dim3(64), 0, ctx->stream, m->theta.as<float>(),
                       m->w2bf.as<uint16_t>(), m->coup.as<double>(), rows, m->out.as<float>(), m->H1.as<float>(),
                       m->H1q.as<uint16_t>(), m->H2.as<float>(), store ? 1 : 0);
    QMB_HIP(hipGetLastError());
}

// alpha/beta/amax of one state set for the coupling whose table starts at tab.
struct StateWork {
    DevBuf alpha, beta, part, amax, stats;
    int64_t cap = 0;
    int nb = 0;
    void ensure(int64_t n) {
        if (n <= cap) return;
        alpha.alloc(n * sizeof(double));
        beta.alloc(n * sizeof(double));
        nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
        part.alloc((size_t)nb * 4 * sizeof(double));
        amax.alloc(sizeof(double));
        stats.alloc(8 * sizeof(double));
        cap = n;
    }
};

void state_logpsi(qmb_ctx *ctx, const qmb_basis *b, const MlpGeom &gm, const float *tab, StateWork &w) {
    w.ensure(b->n_rows);
    hipLaunchKernelGGL(mlp_state_logpsi, dim3(w.nb), dim3(256), 0, ctx->stream, b->n_rows, gm.n_win,
                       gm.win.as<uint8_t>(), gm.sub_a, b->reps.as<uint64_t>(), tab, w.alpha.as<double>(),
                       w.beta.as<double>(), w.part.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_finish_max, dim3(1), dim3(256), 0, ctx->stream, w.part.as<double>(), w.nb,
                       w.amax.as<double>());
    QMB_HIP(hipGetLastError());
}

// Sample table and work arrays of the batched kernels: each sample keeps the
// block count of the single-sample kernels (min(1024, ceil(n/256))).
void ensure_batch(qmb_mlp *m) {
    MlpBatch &B = m->batch;
    if (B.valid) return;
    const int ns = (int)m->samples.size();
    std::vector<BSample> hs(ns);
    std::vector<int> task;
    int64_t off = 0;
    for (int i = 0; i < ns; ++i) {
        const MlpSample &sm = m->samples[i];
        const MlpGeom &gm = m->geoms[sm.geom];
        const qmb_basis *b = sm.basis;
        BSample &d = hs[i];
        d.n = b->n_rows;
        d.off = off;
        d.reps = b->reps.as<uint64_t>();
        d.stab = b->stab.as<uint8_t>();
        d.u = sm.u.as<double>();
        d.win = gm.win.as<uint8_t>();
        d.sub_a = gm.sub_a;
        d.n_win = gm.n_win;
        d.G = b->group->G;
        d.blk0 = (int)task.size();
        d.nblk = (int)std::min<int64_t>(1024, (b->n_rows + 255) / 256);
        for (int k = 0; k < d.nblk; ++k) task.push_back(i);
        off += b->n_rows;
    }
    B.ns = ns;
    B.total_blocks = (int)task.size();
    B.total_rows = off;
    B.samples.alloc((size_t)ns * sizeof(BSample));
    B.task.alloc(task.size() * sizeof(int));
    B.alpha.alloc((size_t)off * sizeof(double));
    B.beta.alloc((size_t)off * sizeof(double));
    B.part.alloc((size_t)B.total_blocks * 4 * sizeof(double));
    B.stats.alloc((size_t)ns * 8 * sizeof(double));
    B.hist.alloc((size_t)ns * 512 * sizeof(unsigned long long));
    B.loss.alloc(sizeof(double));
    QMB_HIP(hipMemcpy(B.samples.p, hs.data(), hs.size() * sizeof(BSample), hipMemcpyHostToDevice));
    QMB_HIP(hipMemcpy(B.task.p, task.data(), task.size() * sizeof(int), hipMemcpyHostToDevice));
    B.valid = true;
}

// Loss (and the dout rows of every sample when want_grad) of all samples, one
// launch per phase; the mean loss is left in batch.loss.
void batch_loss(qmb_mlp *m, bool want_grad) {
    qmb_ctx *ctx = m->ctx;
    ensure_batch(m);
    MlpBatch &B = m->batch;
    const BSample *S = B.samples.as<BSample>();
    const int *task = B.task.as<int>();
    const unsigned nb = (unsigned)B.total_blocks, ns = (unsigned)B.ns;
    hipLaunchKernelGGL(mlp_batch_logpsi, dim3(nb), dim3(256), 0, ctx->stream, S, task, m->out.as<float>(),
                       B.alpha.as<double>(), B.beta.as<double>(), B.part.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_batch_finish_max, dim3(ns), dim3(256), 0, ctx->stream, S, B.part.as<double>(),
                       B.stats.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_batch_overlap, dim3(nb), dim3(256), 0, ctx->stream, S, task, B.alpha.as<double>(),
                       B.beta.as<double>(), B.stats.as<double>(), B.part.as<double>(), B.total_blocks);
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_batch_finish_overlap, dim3(ns), dim3(256), 0, ctx->stream, S, B.part.as<double>(),
                       B.total_blocks, B.stats.as<double>());
    QMB_HIP(hipGetLastError());
    if (!want_grad) {
        hipLaunchKernelGGL(mlp_batch_loss, dim3(1), dim3(64), 0, ctx->stream, B.stats.as<double>(), B.ns,
                           B.loss.as<double>());
        QMB_HIP(hipGetLastError());
        return;
    }
    const int64_t nh = (int64_t)B.ns * 512;
    hipLaunchKernelGGL(mlp_zero_u64, dim3((unsigned)std::min<int64_t>(1024, (nh + 255) / 256)), dim3(256), 0,
                       ctx->stream, B.hist.as<unsigned long long>(), nh);
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_batch_grad, dim3(nb), dim3(256), 0, ctx->stream, S, task, B.alpha.as<double>(),
                       B.beta.as<double>(), B.stats.as<double>(), B.hist.as<unsigned long long>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_batch_finish, dim3((unsigned)std::min<int64_t>(1024, (nh + 255) / 256)), dim3(256), 0,
                       ctx->stream, B.hist.as<unsigned long long>(), B.ns, m->dout.as<float>(), 1.0f / B.ns,
                       B.stats.as<double>(), B.loss.as<double>());
    QMB_HIP(hipGetLastError());
}

// Learned start vector of a k=0 ring basis into dx (device, n_rows values),
// entirely on the device: network rows for (delta, j2), log-amplitude and
// phase per representative, reference phase of the first largest amplitude,
// x_r = |c_r| cos(beta_r - beta_ref); w.stats[0] = ||x||^2 and w.stats[1] =
// ||xi||^2 of the seeded noise. Returns false when ||x||^2 is zero or not
// finite (the caller then starts from the seeded random vector); only those
// two numbers travel to the host.
bool build_start(qmb_mlp *m, const qmb_basis *b, double delta, double j2, uint64_t seed, double *dx,
                 StateWork &w) {
    const int gi = geom_for(m, b);
    qmb_ctx *ctx = m->ctx;
    std::vector<double> coup = {delta, j2};
    forward_rows(m, coup, false);
    state_logpsi(ctx, b, m->geoms[gi], m->out.as<float>(), w);
    DevBuf bref(sizeof(double));
    hipLaunchKernelGGL(mlp_argmax_part, dim3(w.nb), dim3(256), 0, ctx->stream, b->n_rows, w.alpha.as<double>(),
                       w.amax.as<double>(), w.part.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_beta_ref, dim3(1), dim3(256), 0, ctx->stream, w.nb, w.part.as<double>(),
                       w.beta.as<double>(), bref.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_state_start, dim3(w.nb), dim3(256), 0, ctx->stream, b->n_rows, b->group->G,
                       b->stab.as<uint8_t>(), w.alpha.as<double>(), w.beta.as<double>(), w.amax.as<double>(),
                       bref.as<double>(), seed, dx, w.part.as<double>());
    QMB_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_finish_sum2, dim3(1), dim3(256), 0, ctx->stream, w.part.as<double>(), w.nb,
                       w.stats.as<double>());
    QMB_HIP(hipGetLastError());
    double sums[2] = {0.0, 0.0};
    QMB_HIP(hipMemcpyAsync(sums, w.stats.p, sizeof(sums), hipMemcpyDeviceToHost, ctx->stream));
    QMB_HIP(hipStreamSynchronize(ctx->stream));
    return std::isfinite(sums[0]) && sums[0] > 0.0 && sums[1] > 0.0;
}

// Mean loss over the samples and, optionally, the gradient (device grad buffer).
double loss_and_grad(qmb_mlp *m, bool want_grad) {
    qmb_ctx *ctx = m->ctx;
    const int ns = (int)m->samples.size();
    std::vector<double> coup(2 * ns);
    for (int i = 0; i < ns; ++i) {
        coup[2 * i] = m->samples[i].delta;
        coup[2 * i + 1] = m->samples[i].j2;
    }
    forward_rows(m, coup, want_grad);
    const int rows = ns * 256;
    batch_loss(m, want_grad);
    if (want_grad) {
        hipLaunchKernelGGL(mlp_backward_act, dim3((rows + 31) / 32), dim3(64), 0, ctx->stream, m->theta.as<float>(),
                           m->w2t.as<uint16_t>(), rows, m->dout.as<float>(), m->H1.as<float>(), m->H2.as<float>(),
                           m->dA2.as<float>(), m->dA2q.as<uint16_t>(), m->dA1.as<float>());
        QMB_HIP(hipGetLastError());
        const int groups = (rows + kRowsPerGroup - 1) / kRowsPerGroup;
        hipLaunchKernelGGL(mlp_wgrad2, dim3(groups, 16), dim3(64), 0, ctx->stream, rows, m->dA2q.as<uint16_t>(),
                           m->H1q.as<uint16_t>(), m->part2.as<float>());
        QMB_HIP(hipGetLastError());
        const int groups_s = (rows + kRowsPerSmallGroup - 1) / kRowsPerSmallGroup;
        hipLaunchKernelGGL(mlp_wgrad_small, dim3(groups_s), dim3(kH), 0, ctx->stream, m->coup.as<double>(), rows,
                           m->dout.as<float>(), m->H2.as<float>(), m->dA2.as<float>(), m->dA1.as<float>(),
                           m->part.as<float>());
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(mlp_reduce_grad, dim3((kParams + 255) / 256), dim3(256), 0, ctx->stream, groups_s, groups,
                           m->part.as<float>(), m->part2.as<float>(), 1.0f, m->grad.as<float>());
        QMB_HIP(hipGetLastError());
    }
    double *hl = ctx->host_stage<double>(sizeof(double));
    QMB_HIP(hipMemcpyAsync(hl, m->batch.loss.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    QMB_HIP(hipStreamSynchronize(ctx->stream));
    return *hl;
}

std::vector<float> init_params(uint64_t seed) {
    std::vector<float> th(kParams);
    const int fan_for[] = {kIn, kIn, kH, kH, kH, kH};
    const int ends[] = {ob1, oW2, ob2, oW3, ob3, kParams};
    int seg = 0;
    for (int i = 0; i < kParams; ++i) {
        while (i >= ends[seg]) ++seg;
        uint64_t z = seed + ((uint64_t)i + 1ULL) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        const double u = (double)((z >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
        th[i] = (float)((2.0 * u - 1.0) / std::sqrt((double)fan_for[seg]));
    }
    return th;
}

}  // namespace

extern "C" {

int qmb_mlp_num_params(void) { return kParams; }

qmb_status qmb_mlp_create(qmb_ctx *ctx, uint64_t seed, qmb_mlp **out) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && out, "qmb_mlp_create: NULL argument");
        *out = nullptr;
        QMB_HIP(hipSetDevice(ctx->device));
        auto m = std::make_unique<qmb_mlp>();
        m->ctx = ctx;
        m->theta.alloc(kParams * sizeof(float));
        m->m1.alloc(kParams * sizeof(float));
        m->m2.alloc(kParams * sizeof(float));
        m->grad.alloc(kParams * sizeof(float));
        m->w2bf.alloc(kH * kH * sizeof(uint16_t));
        m->w2t.alloc(kH * kH * sizeof(uint16_t));
        const std::vector<float> th = init_params(seed);
        QMB_HIP(hipMemcpy(m->theta.p, th.data(), kParams * sizeof(float), hipMemcpyHostToDevice));
        QMB_HIP(hipMemset(m->m1.p, 0, kParams * sizeof(float)));
        QMB_HIP(hipMemset(m->m2.p, 0, kParams * sizeof(float)));
        hipLaunchKernelGGL(mlp_refresh_w2, dim3(kH * kH / 256), dim3(256), 0, ctx->stream, m->theta.as<float>(),
                           m->w2bf.as<uint16_t>(), m->w2t.as<uint16_t>());
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        *out = m.release();
    });
}

qmb_status qmb_mlp_destroy(qmb_mlp *m) {
    if (m && m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    release(m);
    return QMB_OK;
}

qmb_status qmb_mlp_get_params(const qmb_mlp *m, float *theta) {
    return guarded([&] {
        QMB_REQUIRE(m && theta, "qmb_mlp_get_params: NULL argument");
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
        QMB_HIP(hipMemcpy(theta, m->theta.p, kParams * sizeof(float), hipMemcpyDeviceToHost));
    });
}

qmb_status qmb_mlp_set_params(qmb_mlp *m, const float *theta) {
    return guarded([&] {
        QMB_REQUIRE(m && theta, "qmb_mlp_set_params: NULL argument");
        QMB_HIP(hipMemcpy(m->theta.p, theta, kParams * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(mlp_refresh_w2, dim3(kH * kH / 256), dim3(256), 0, m->ctx->stream, m->theta.as<float>(),
                           m->w2bf.as<uint16_t>(), m->w2t.as<uint16_t>());
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
    });
}

qmb_status qmb_mlp_forward(qmb_mlp *m, int n, const double *delta, const double *j2, float *out) {
    return guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && delta && j2 && out && n >= 1, "qmb_mlp_forward: bad argument");
        std::vector<double> coup(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            coup[2 * i] = delta[i];
            coup[2 * i + 1] = j2[i];
        }
        forward_rows(m, coup, false);
        QMB_HIP(hipMemcpyAsync(out, m->out.p, (size_t)n * 512 * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
    });
}

qmb_status qmb_mlp_add_sample(qmb_mlp *m, const qmb_basis *b, const double *ground_state, double delta, double j2) {
    return guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && b && ground_state, "qmb_mlp_add_sample: NULL argument");
        QMB_REQUIRE(b->ctx == m->ctx, "qmb_mlp_add_sample: basis from another context");
        MlpSample s;
        s.geom = geom_for(m, b);
        s.basis = b;
        s.delta = delta;
        s.j2 = j2;
        s.u.alloc((size_t)b->n_rows * sizeof(double));
        QMB_HIP(hipMemcpy(s.u.p, ground_state, (size_t)b->n_rows * sizeof(double), hipMemcpyHostToDevice));
        m->samples.push_back(std::move(s));
        m->batch.valid = false;
    });
}

qmb_status qmb_mlp_set_windows(qmb_mlp *m, const qmb_basis *b, int n_win, const uint8_t *sites,
                               uint64_t sublattice_a) {
    return guarded([&] {
        QMB_REQUIRE(m && b && sites, "qmb_mlp_set_windows: NULL argument");
        QMB_REQUIRE(b->ctx == m->ctx, "qmb_mlp_set_windows: basis from another context");
        check_mlp_basis(b);
        QMB_REQUIRE(n_win >= 1 && n_win <= kMaxWin, "qmb_mlp_set_windows: n_win must be in [1, 128]");
        for (int i = 0; i < n_win * kWin; ++i)
            QMB_REQUIRE(sites[i] < b->n_sites, "qmb_mlp_set_windows: site index out of range");
        const uint64_t full = b->n_sites == 64 ? ~0ULL : ((1ULL << b->n_sites) - 1ULL);
        QMB_REQUIRE((sublattice_a & ~full) == 0, "qmb_mlp_set_windows: sublattice mask beyond the lattice");
        for (const auto &s : m->samples)
            QMB_REQUIRE(s.basis != b, "qmb_mlp_set_windows: the basis already has samples");
        MlpGeom gm;
        gm.basis = b;
        gm.n_win = n_win;
        gm.win.alloc((size_t)n_win * kWin);
        QMB_HIP(hipMemcpy(gm.win.p, sites, (size_t)n_win * kWin, hipMemcpyHostToDevice));
        gm.sub_a = sublattice_a;
        for (auto &g : m->geoms)
            if (g.basis == b) {
                g = std::move(gm);
                return;
            }
        m->geoms.push_back(std::move(gm));
    });
}

qmb_status qmb_mlp_clear_samples(qmb_mlp *m) {
    return guarded([&] {
        QMB_REQUIRE(m, "qmb_mlp_clear_samples: NULL argument");
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
        m->samples.clear();
        m->batch = MlpBatch{};
    });
}

qmb_status qmb_mlp_loss_grad(qmb_mlp *m, double *loss, float *grad) {
    return guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && loss, "qmb_mlp_loss_grad: NULL argument");
        QMB_REQUIRE(!m->samples.empty(), "qmb_mlp_loss_grad: no samples");
        *loss = loss_and_grad(m, grad != nullptr);
        if (grad) QMB_HIP(hipMemcpy(grad, m->grad.p, kParams * sizeof(float), hipMemcpyDeviceToHost));
    });
}

qmb_status qmb_mlp_train(qmb_mlp *m, int steps, double lr, double *loss_history) {
    return guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && steps >= 1 && lr > 0, "qmb_mlp_train: bad argument");
        QMB_REQUIRE(!m->samples.empty(), "qmb_mlp_train: no samples");
        const float b1 = 0.9f, b2 = 0.999f;
        for (int s = 0; s < steps; ++s) {
            const double L = loss_and_grad(m, true);
            if (loss_history) loss_history[s] = L;
            ++m->step;
            const float bc1 = 1.f - std::pow(b1, (float)m->step), bc2 = 1.f - std::pow(b2, (float)m->step);
            hipLaunchKernelGGL(mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, m->ctx->stream,
                               m->theta.as<float>(), m->grad.as<float>(), m->m1.as<float>(), m->m2.as<float>(),
                               (float)lr, b1, b2, bc1, bc2, m->w2bf.as<uint16_t>(), m->w2t.as<uint16_t>());
            QMB_HIP(hipGetLastError());
        }
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
    });
}

qmb_status qmb_mlp_start_vector(qmb_mlp *m, const qmb_basis *b, double delta, double j2, double *x, int *fallback) {
    return guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && b && x, "qmb_mlp_start_vector: NULL argument");
        DevBuf dx((size_t)b->n_rows * sizeof(double));
        StateWork w;
        const bool ok = build_start(m, b, delta, j2, 0, dx.as<double>(), w);
        QMB_HIP(hipMemcpyAsync(x, dx.p, b->n_rows * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        QMB_HIP(hipStreamSynchronize(m->ctx->stream));
        if (fallback) *fallback = ok ? 0 : 1;
    });
}

qmb_status qmb_lanczos_learned(qmb_mlp *m, const qmb_basis *b, const qmb_ell *e, double delta, double j2,
                               const qmb_lanczos_opts *opts, double *evals, double *eigvec, qmb_lanczos_report *report,
                               int *fallback) {
    qmb_status st = guarded([&] {
        PoolScope pool_scope(m ? &m->ctx->pool : nullptr);
        QMB_REQUIRE(m && b && e && evals, "qmb_lanczos_learned: NULL argument");
        QMB_REQUIRE(e->basis == b, "qmb_lanczos_learned: the ELL was not built on this basis");
        QMB_REQUIRE(e->ctx == m->ctx, "qmb_lanczos_learned: network and ELL live on different contexts");
        QMB_REQUIRE(e->value_type == QMB_REAL, "qmb_lanczos_learned: the learned start vector is real; "
                                               "build the ELL with QMB_REAL");
    });
    if (st != QMB_OK) return st;
    qmb_lanczos_opts o;
    if (opts)
        o = *opts;
    else
        qmb_lanczos_default_opts(&o);
    DevBuf dx;
    StateWork w;
    bool ok = false;
    st = guarded([&] {
        dx.alloc((size_t)b->n_rows * sizeof(double));
        ok = build_start(m, b, delta, j2, o.seed, dx.as<double>(), w);
        if (ok) {
            // unit learned vector + kStartNoise x unit seeded noise: the noise keeps
            // every eigenvector in the Krylov space even where the network's
            // vector misses the ground state
            hipLaunchKernelGGL(mlp_mix_noise, dim3(w.nb), dim3(256), 0, m->ctx->stream, b->n_rows, o.seed,
                               kStartNoise, w.stats.as<double>(), dx.as<double>());
            QMB_HIP(hipGetLastError());
        }
    });
    if (st != QMB_OK) return st;
    if (fallback) *fallback = ok ? 0 : 1;
    if (ok) {
        o.start = QMB_START_DEVICE;
        o.start_vec = dx.as<double>();
    } else {
        o.start = QMB_START_RANDOM;
        o.start_vec = nullptr;
    }
    st = qmb_lanczos(m->ctx, e, &o, evals, nullptr, eigvec, report);
    if (ok && st == QMB_ERR_NOT_CONVERGED) {
        // the Ritz residual of the learned run missed the tolerance within max_iter:
        // the seeded random start is the fallback and its run is the result
        if (fallback) *fallback = 2;
        o.start = QMB_START_RANDOM;
        o.start_vec = nullptr;
        st = qmb_lanczos(m->ctx, e, &o, evals, nullptr, eigvec, report);
    }
    return st;
}

}  // extern "C"
