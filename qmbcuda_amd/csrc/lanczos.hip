// ELL SpMV and the FP64 Lanczos loop.
//
// Replaces SparseMatDenseVec_prod_EEL (Q/exact_diagonalization/utils.cu:601-624)
// and LanczosEigVals_CUDA / LanczosEigenStateFromSeed (Q/exact_diagonalization/
// lanczos.cu:31-317). The recurrence is the reference's, in FP64:
//   b[0] = 0; phi_0 = start / ||start||;
//   y = H phi_n; a[n] = Re<phi_n, y>; w = (-a[n]) phi_n + (y + (-b[n]) phi_{n-1});
//   b[n+1] = ||w||; phi_{n+1} = w * (1/b[n+1]).
// Per step there are two streaming kernels and no host round trip:
//   K1 spmv_sell<LZ>: y = (1/b) (H w) (the scale applied once per row after
//      the gathers), phi_n = w/b written out, the block partials of a[n] in
//      the same pass;
//   K2 lz_update: w = (-a) phi_n + (y - b phi_{n-1}) and the partials of ||w||^2.
// In the plain recurrence every block of K2 sums K1's partials into a[n] and
// every block of the next K1 sums K2's into b[n+1] (same fixed order as a
// separate one-block finish_sum, which runs only before a convergence check and
// with reorthogonalisation). The host reads a[], b[] every check_every steps and solves the
// tridiagonal problem in FP64 (lanczos_host.cpp). Reductions are deterministic:
// fixed block partial slots summed in a fixed order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>

#include "checkpoint.hpp"
#include "internal.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace qmb {

constexpr int kMaxBlocks = 4096;

// ---- K1: SELL-64 SpMV --------------------------------------------------------
// One wavefront per 64-row chunk (row products in sell.hpp); the x gathers are
// the only random accesses. LZ fuses the Lanczos normalisation and the a[n]
// dot: y = s (H x), phi = s x, partial sums of phi . y. PACK: the packed value
// format (4 B per off-diagonal entry, value tables staged in LDS; complex: real
// parts at [0, n_values), imaginary parts at kMaxValues + [0, n_values_im)).
template <bool CPLX, bool LZ, bool PACK, bool NT>
__global__ __launch_bounds__(kThreads) void spmv_sell(int64_t n_rows, int64_t n_chunks,
                                                      const int64_t *__restrict__ chunk_ptr,
                                                      const uint16_t *__restrict__ row_len,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ vtab, int n_values, int cb,
                                                      int n_values_im, int rb,
                                                      const double *__restrict__ x, double *__restrict__ y,
                                                      const double *__restrict__ invb, double *__restrict__ vout,
                                                      double *__restrict__ partials, const double *__restrict__ bpart,
                                                      int n_bpart, double *__restrict__ beta_out,
                                                      double *__restrict__ invb_out) {
    __shared__ double lds[kWaves];
    __shared__ double s_vt[PACK ? (CPLX ? 2 : 1) * kMaxValues : 1];
    __shared__ double s_scale;
    if (PACK) {
        for (int i = threadIdx.x; i < n_values; i += blockDim.x) s_vt[i] = vtab[i];
        if (CPLX)
            for (int i = threadIdx.x; i < n_values_im; i += blockDim.x) s_vt[kMaxValues + i] = vtab[kMaxValues + i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    // Fused b[n]: the previous step's ||w||^2 partials, summed by every block in
    // the fixed order of finish_sum (thread t adds partials t, t + blockDim, ...
    // in order, then the block sum): b = sqrt(sum), s = 1/b (0 if b = 0). The
    // first kPre partials of each thread are loaded at once, and s is formed
    // after the block's first chunk product (which does not need it), so the
    // sum's latency overlaps the SpMV.
    constexpr int kPre = 4;
    const bool fused = LZ && bpart != nullptr;
    double pv[kPre];
    if (fused) {
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int i = (int)threadIdx.x + u * (int)blockDim.x;
            pv[u] = i < n_bpart ? bpart[i] : 0.0;
        }
    }
    double s = (LZ && !fused) ? *invb : 1.0;
    auto form_scale = [&]() {
        double t = 0.0;
#pragma unroll
        for (int u = 0; u < kPre; ++u)
            if ((int)threadIdx.x + u * (int)blockDim.x < n_bpart) t += pv[u];
        for (int i = (int)threadIdx.x + kPre * (int)blockDim.x; i < n_bpart; i += blockDim.x)
            t += bpart[i];
        t = block_sum(t, lds);
        if (threadIdx.x == 0) {
            const double nb = sqrt(t);
            const double inv = nb > 0.0 ? 1.0 / nb : 0.0;
            s_scale = inv;
            if (blockIdx.x == 0) {
                *beta_out = nb;
                *invb_out = inv;
            }
        }
        __syncthreads();
        s = s_scale;
    };
    double dot = 0.0;
    // row product of chunk c (unscaled) and its output with the scale s
    auto product = [&](int64_t c, double2 &acc, double2 &xr) {
        const int64_t r = c * 64 + lane;
        const int64_t base = chunk_ptr[c];
        const int L = (int)((chunk_ptr[c + 1] - base) >> 6);
        const int len = r < n_rows ? (int)row_len[r] : 0;
        const int64_t p0 = base + lane;
        if (!CPLX) {
            if (PACK) {
                xr.x = r < n_rows ? x[r] : 0.0;
                const double d = r < n_rows ? diag[r] : 0.0;
                const uint32_t *wd = reinterpret_cast<const uint32_t *>(col);
                acc.x = sell_row_packed<NT>(wd, s_vt, x, p0, L, len, cb, d, xr.x);
            } else {
                acc.x = sell_row_real<NT>(col, val, x, p0, L, len);
                if (LZ) xr.x = r < n_rows ? x[r] : 0.0;
            }
        } else {
            xr = r < n_rows ? reinterpret_cast<const double2 *>(x)[r] : make_double2(0.0, 0.0);
            if (PACK) {
                const double2 d = r < n_rows ? reinterpret_cast<const double2 *>(diag)[r] : make_double2(0.0, 0.0);
                const uint32_t *wd = reinterpret_cast<const uint32_t *>(col);
                acc = sell_row_cplx_packed<NT>(wd, s_vt, s_vt + kMaxValues, x, p0, L, len, cb, rb, d, xr);
            } else {
                acc = sell_row_cplx<NT>(col, val, x, p0, L, len);
            }
        }
    };
    auto emit = [&](int64_t c, double2 acc, double2 xr) {
        const int64_t r = c * 64 + lane;
        if (!CPLX) {
            double a = acc.x;
            if (LZ) a *= s;
            if (r < n_rows) {
                y[r] = a;
                if (LZ) {
                    const double vr = xr.x * s;
                    vout[r] = vr;
                    dot += vr * a;
                }
            }
        } else {
            double2 a = acc;
            if (LZ) {
                a.x *= s;
                a.y *= s;
            }
            if (r < n_rows) {
                reinterpret_cast<double2 *>(y)[r] = a;
                if (LZ) {
                    double2 vr = xr;
                    vr.x = vr.x * s;
                    vr.y = vr.y * s;
                    reinterpret_cast<double2 *>(vout)[r] = vr;
                    dot += vr.x * a.x + vr.y * a.y;  // Re(conj(v) y)
                }
            }
        }
    };
    // Streamed (large) matrices: XCD-contiguous chunk assignment. The hardware
    // places block b on XCD b % 8; logical block lb gives each XCD one contiguous
    // eighth of every pass over the chunks, so x lines that nearby rows gather stay
    // in that XCD's L2 (on the 6x6 sector 37 % of the targets lie within 16K rows
    // of their row: K1 27.6 -> 26.1 ms). A cached matrix keeps the interleaved
    // assignment (the N=28 chain's K1 went from 51 to 60 us with it).
    const unsigned nb = gridDim.x;
    const unsigned lb = (NT && nb % 8u == 0u) ? (blockIdx.x % 8u) * (nb / 8u) + blockIdx.x / 8u : blockIdx.x;
    const int64_t cstride = (int64_t)gridDim.x * kWaves;
    int64_t c = (int64_t)lb * kWaves + wv;
    if (fused) {
        // every thread joins the block sum, with or without a chunk
        double2 acc0 = make_double2(0.0, 0.0), xr0 = make_double2(0.0, 0.0);
        const bool have = c < n_chunks;
        if (have) product(c, acc0, xr0);
        form_scale();
        if (have) emit(c, acc0, xr0);
        c += cstride;
    }
    for (; c < n_chunks; c += cstride) {
        double2 acc = make_double2(0.0, 0.0), xr = make_double2(0.0, 0.0);
        product(c, acc, xr);
        emit(c, acc, xr);
    }
    if (LZ) {
        const double t = block_sum(dot, lds);
        if (threadIdx.x == 0) partials[blockIdx.x] = t;
    }
}

// Second Gram-Schmidt pass only when the first one removed more than this
// share of the norm (Daniel, Gragg, Kaufman, Stewart, Math. Comp. 30 (1976)):
// one pass is then already orthogonal to working precision.
constexpr double kDgks = 0.7071067811865476;

// gate: NULL, or a device flag; kernels of a gated second pass return at once
// when it is 0.
__device__ inline bool gated_off(const double *gate) { return gate && *gate == 0.0; }

// Fixed-order sum of the block partials of the previous kernel (one block):
// mode 0 -> out = s; mode 1 -> out = sqrt(s), out_inv = 1/sqrt(s) (0 if s = 0),
// skipped when the gate is 0; mode 2 (after a first Gram-Schmidt pass) -> as
// mode 1, and *gate = 1 if sqrt(s) < kDgks * (the previous *out), else 0.
// A separate launch instead of a "last block" finish: on this chip a
// device-scope fence in every block writes back and invalidates the L2 and
// costs more than the SpMV's own traffic.
__global__ __launch_bounds__(kThreads) void finish_sum(const double *__restrict__ partials, int n,
                                                       double *__restrict__ out, double *__restrict__ out_inv,
                                                       int mode, double *__restrict__ gate) {
    __shared__ double lds[kWaves];
    if (mode == 1 && gated_off(gate)) return;
    const double s = sum_partials(partials, n, lds);
    if (threadIdx.x == 0) {
        if (mode == 0) {
            *out = s;
        } else {
            const double nb = sqrt(s);
            if (mode == 2) *gate = nb < kDgks * *out ? 1.0 : 0.0;
            if (out) *out = nb;
            if (out_inv) *out_inv = nb > 0.0 ? 1.0 / nb : 0.0;
        }
    }
}

// ---- K2: three-term update + ||w||^2 ----------------------------------------
__global__ __launch_bounds__(kThreads) void lz_update(int64_t cnt, const double *__restrict__ y,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ vprev, double *__restrict__ w,
                                                      double *__restrict__ alpha_p,
                                                      const double *__restrict__ beta_p,
                                                      double *__restrict__ partials,
                                                      const double *__restrict__ apart, int n_apart) {
    __shared__ double lds[kWaves];
    __shared__ double s_alpha;
    // The thread's first U elements are loaded before the a[n] sum, so their
    // latency overlaps it; the sum over the elements keeps index order.
    constexpr int U = 4;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double ry[U], rv[U], rp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * st;
        ry[u] = i < cnt ? y[i] : 0.0;
        rv[u] = i < cnt ? v[i] : 0.0;
        rp[u] = i < cnt ? vprev[i] : 0.0;
    }
    double a;
    if (apart) {
        // a[n] from the SpMV's partials, summed by every block in finish_sum's order
        const double t = sum_partials(apart, n_apart, lds);
        if (threadIdx.x == 0) {
            s_alpha = t;
            if (blockIdx.x == 0) *alpha_p = t;
        }
        __syncthreads();
        a = s_alpha;
    } else {
        a = *alpha_p;
    }
    const double b = *beta_p;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * st;
        if (i < cnt) {
            const double t = ry[u] + (-b) * rp[u];
            const double wi = (-a) * rv[u] + t;
            w[i] = wi;
            acc += wi * wi;
        }
    }
    for (int64_t i = i0 + U * st; i < cnt; i += st) {
        const double t = y[i] + (-b) * vprev[i];
        const double wi = (-a) * v[i] + t;
        w[i] = wi;
        acc += wi * wi;
    }
    const double tb = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = tb;
}

// ---- start vector, norms, eigenvector helpers -------------------------------
__global__ __launch_bounds__(kThreads) void lz_random(int64_t cnt, uint64_t seed, double *w) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = splitmix_uniform(seed, (uint64_t)i);
}

// out_norm = ||a - e*b|| (b may be NULL -> ||a||); out_inv = 1/out_norm.
__global__ __launch_bounds__(kThreads) void vec_norm(int64_t cnt, const double *__restrict__ a,
                                                     const double *__restrict__ b, double e,
                                                     double *__restrict__ partials) {
    __shared__ double lds[kWaves];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        const double d = b ? a[i] - e * b[i] : a[i];
        acc += d * d;
    }
    const double t = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// x = x + c * v (the reference's saxpy_functor2 accumulation of the eigenvector).
__global__ __launch_bounds__(kThreads) void vec_axpy(int64_t cnt, double c, const double *__restrict__ v,
                                                     double *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x[i] + c * v[i];
}

// ---- reorthogonalisation (classical Gram-Schmidt, second pass on demand) ------
// dots: coef[j] = <V_j, w> for j < nv; grid (blocks, ceil(nv / kDotGroup)). A
// block takes kDotGroup vectors at once, so w is read once per group instead of
// once per vector; every dot keeps its own accumulator in the same index order,
// so the partials are those of one vector per block.
constexpr int kDotGroup = 4;

template <bool CPLX>
__global__ __launch_bounds__(kThreads) void reorth_dots(int64_t n, int64_t stride, int nv,
                                                        const double *__restrict__ V,
                                                        const double *__restrict__ w, double *__restrict__ part,
                                                        int nb, const double *__restrict__ gate) {
    __shared__ double lds[kWaves];
    if (gated_off(gate)) return;
    const int j0 = blockIdx.y * kDotGroup;
    const int nj = min(kDotGroup, nv - j0);
    double sr[kDotGroup], si[kDotGroup];
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) {
        sr[q] = 0.0;
        si[q] = 0.0;
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            const double wr = w[2 * i], wi = w[2 * i + 1];
#pragma unroll
            for (int q = 0; q < kDotGroup; ++q) {
                if (q < nj) {
                    const double *vj = V + (size_t)(j0 + q) * stride;
                    const double vr = vj[2 * i], vi = vj[2 * i + 1];
                    sr[q] += vr * wr + vi * wi;
                    si[q] += vr * wi - vi * wr;
                }
            }
        } else {
            const double wr = w[i];
#pragma unroll
            for (int q = 0; q < kDotGroup; ++q)
                if (q < nj) sr[q] += V[(size_t)(j0 + q) * stride + i] * wr;
        }
    }
#pragma unroll
    for (int q = 0; q < kDotGroup; ++q) {
        if (q < nj) {  // nj is uniform over the block: every thread takes part in each sum
            const double tr = block_sum(sr[q], lds);
            const double ti = CPLX ? block_sum(si[q], lds) : 0.0;
            if (threadIdx.x == 0) {
                part[((size_t)(j0 + q) * nb + blockIdx.x) * 2] = tr;
                part[((size_t)(j0 + q) * nb + blockIdx.x) * 2 + 1] = ti;
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void reorth_reduce(const double *__restrict__ part, int nb,
                                                          double *__restrict__ coef, const double *__restrict__ gate) {
    __shared__ double lds[kWaves];
    if (gated_off(gate)) return;
    const int j = blockIdx.x;
    double sr = 0.0, si = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        sr += part[((size_t)j * nb + i) * 2];
        si += part[((size_t)j * nb + i) * 2 + 1];
    }
    const double tr = block_sum(sr, lds);
    const double ti = block_sum(si, lds);
    if (threadIdx.x == 0) {
        coef[2 * j] = tr;
        coef[2 * j + 1] = ti;
    }
}

// w -= sum_j coef_j V_j; block partials of ||w||^2.
template <bool CPLX>
__global__ __launch_bounds__(kThreads) void reorth_apply(int64_t n, int64_t stride, int nv,
                                                         const double *__restrict__ V,
                                                         const double *__restrict__ coef, double *__restrict__ w,
                                                         double *__restrict__ partials,
                                                         const double *__restrict__ gate) {
    __shared__ double lds[kWaves];
    if (gated_off(gate)) return;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (CPLX) {
            double wr = w[2 * i], wi = w[2 * i + 1];
            for (int j = 0; j < nv; ++j) {
                const double cr = coef[2 * j], ci = coef[2 * j + 1];
                const double vr = V[(size_t)j * stride + 2 * i], vi = V[(size_t)j * stride + 2 * i + 1];
                wr -= cr * vr - ci * vi;
                wi -= cr * vi + ci * vr;
            }
            w[2 * i] = wr;
            w[2 * i + 1] = wi;
            acc += wr * wr + wi * wi;
        } else {
            double wr = w[i];
            for (int j = 0; j < nv; ++j) wr -= coef[2 * j] * V[(size_t)j * stride + i];
            w[i] = wr;
            acc += wr * wr;
        }
    }
    const double t = block_sum(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// x = sum_j z_j V_j (eigenvector from a stored basis).
__global__ __launch_bounds__(kThreads) void basis_combine(int64_t cnt, int64_t stride, int nv,
                                                          const double *__restrict__ V, const double *__restrict__ z,
                                                          double *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < nv; ++j) s = s + z[j] * V[(size_t)j * stride + i];
        x[i] = s;
    }
}

// ---- launch helpers -----------------------------------------------------------
static unsigned grid_for(int64_t work_items, int per_block) {
    const int64_t b = (work_items + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, kMaxBlocks));
}

struct LzWork {
    DevBuf partials, scal;  // scal: alpha[M], beta[M+1], invb[M+1], misc[8]
    int M = 0;
    double *alpha() const { return scal.as<double>(); }
    double *beta() const { return scal.as<double>() + M; }
    double *invb() const { return scal.as<double>() + 2 * M + 1; }
    double *misc() const { return scal.as<double>() + 3 * M + 2; }
    void init(qmb_ctx *ctx, int max_iter) {
        M = max_iter;
        partials.alloc(kMaxBlocks * sizeof(double) * 2);
        scal.alloc((size_t)(3 * M + 2 + 8) * sizeof(double));
        QMB_HIP(hipMemsetAsync(scal.p, 0, (size_t)(3 * M + 2 + 8) * sizeof(double), ctx->stream));
    }
};

static void launch_finish(qmb_ctx *ctx, const double *partials, unsigned n, double *out, double *out_inv, int mode,
                          double *gate = nullptr) {
    hipLaunchKernelGGL(finish_sum, dim3(1), dim3(kThreads), 0, ctx->stream, partials, (int)n, out, out_inv, mode,
                       gate);
    QMB_HIP(hipGetLastError());
}

// after (optional): an event recorded between the SpMV and the a[n] finish, so
// that profiling brackets the SpMV kernel alone.
// fused (optional): the previous step's ||w||^2 partials (n_bpart of them) from
// which every block derives 1/b itself, writing b and 1/b to beta_out/invb_out;
// the a[n] finish is then left to lz_update (alpha_out unused).
struct FusedBeta {
    const double *bpart = nullptr;
    int n_bpart = 0;
    double *beta_out = nullptr, *invb_out = nullptr;
    double *partials = nullptr;  // where this SpMV's a[n] partials go
};

// MI355X's memory-side cache (MALL) holds 256 MB: a packed matrix that fits it
// beside the Lanczos vectors is read with plain loads, so its lines stay cached
// from one SpMV to the next; a larger one is streamed (nontemporal loads).
constexpr double kMallBytes = 256.0 * 1024 * 1024;

void launch_spmv(qmb_ctx *ctx, const qmb_ell *e, const double *x, double *y, bool lz, const double *invb,
                 double *vout, LzWork *wk, double *alpha_out, hipEvent_t after = nullptr,
                 const FusedBeta *fb = nullptr) {
    const unsigned blocks = grid_for(e->n_chunks, kWaves);
    const double vec_bytes = (e->value_type == QMB_COMPLEX ? 16.0 : 8.0) * (double)e->n_rows;
    const bool stream = e->format_bytes() + 8.0 * vec_bytes > 0.75 * kMallBytes;
    const bool cplx = e->value_type == QMB_COMPLEX;
    using K = decltype(&spmv_sell<false, false, false, false>);
    // [complex][packed][lz][stream]
    static const K table[2][2][2][2] = {
        {{{spmv_sell<false, false, false, false>, spmv_sell<false, false, false, true>},
          {spmv_sell<false, true, false, false>, spmv_sell<false, true, false, true>}},
         {{spmv_sell<false, false, true, false>, spmv_sell<false, false, true, true>},
          {spmv_sell<false, true, true, false>, spmv_sell<false, true, true, true>}}},
        {{{spmv_sell<true, false, false, false>, spmv_sell<true, false, false, true>},
          {spmv_sell<true, true, false, false>, spmv_sell<true, true, false, true>}},
         {{spmv_sell<true, false, true, false>, spmv_sell<true, false, true, true>},
          {spmv_sell<true, true, true, false>, spmv_sell<true, true, true, true>}}}};
    const K k = table[cplx ? 1 : 0][e->packed ? 1 : 0][lz ? 1 : 0][stream ? 1 : 0];
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), 0, ctx->stream, e->n_rows, e->n_chunks,
                       e->chunk_ptr.as<int64_t>(), e->row_len.as<uint16_t>(), e->col.as<int32_t>(),
                       e->val.as<double>(), e->diag.as<double>(), e->vtab.as<double>(), e->n_values, e->col_bits,
                       e->n_values_im, e->re_bits, x, y, invb, vout, fb ? fb->partials : wk ? wk->partials.as<double>() : nullptr,
                       fb ? fb->bpart : nullptr, fb ? fb->n_bpart : 0, fb ? fb->beta_out : nullptr,
                       fb ? fb->invb_out : nullptr);
    QMB_HIP(hipGetLastError());
    if (after) QMB_HIP(hipEventRecord(after, ctx->stream));
    if (lz && !fb) launch_finish(ctx, wk->partials.as<double>(), blocks, alpha_out, nullptr, 0);
}

unsigned spmv_blocks(const qmb_ell *e) { return grid_for(e->n_chunks, kWaves); }

static void launch_norm(qmb_ctx *ctx, int64_t cnt, const double *a, const double *b, double e, LzWork &wk,
                        double *out_norm, double *out_inv) {
    const unsigned blocks = grid_for(cnt, kThreads * 4);
    hipLaunchKernelGGL(vec_norm, dim3(blocks), dim3(kThreads), 0, ctx->stream, cnt, a, b, e,
                       wk.partials.as<double>());
    QMB_HIP(hipGetLastError());
    launch_finish(ctx, wk.partials.as<double>(), blocks, out_norm, out_inv, 1);
}

}  // namespace qmb

using namespace qmb;

namespace {

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Lanczos checkpoint: the exact state at a convergence check (steps done,
// a[], b[], 1/b[], the next unnormalised vector, the vector(s) the recurrence
// still needs, the orthogonality estimates and the check schedule), so a
// resumed run continues with the same operations and ends with the same
// results as an uninterrupted one.
struct Checkpoint {
    const char *path = nullptr;
    int every = 0;       // write at the first check at least this many steps after the last write
    int resume = 0;      // start from the file instead of the start vector
    int stop_after = 0;  // stop (and write) at the first check at least this many steps into this call
};

struct LanczosFileHeader {
    char magic[8];
    uint32_t version, header_bytes;
    int64_t n_rows, nnz;
    int32_t value_type, reorth, max_iter, n_eigs, start, fixed_iter, check_every, pad0;
    uint64_t seed;
    double tol;
    int32_t it, n_reorth, pro_second, next, it_prev, n_spmv;
    double q_prev;
};
constexpr char kLanczosMagic[8] = {'Q', 'M', 'B', 'L', 'A', 'N', 'C', '1'};

// The Lanczos engine shared by qmb_lanczos and the reference drop-in.
qmb_status run_lanczos(qmb_ctx *ctx, const qmb_ell *e, const qmb_lanczos_opts &o, double *evals, double *krylov,
                       double *eigvec, qmb_lanczos_report *rep, const Checkpoint *ck = nullptr) {
    Timer wall;
    const bool cplx = e->value_type == QMB_COMPLEX;
    const int64_t n = e->n_rows;
    const int64_t cnt = cplx ? 2 * n : n;
    const int M = o.max_iter;
    const int k_eigs = o.n_eigs;
    // reorth 0: plain recurrence; 1: full (every step); 2: partial, driven by
    // the estimated loss of orthogonality; 3: selective, against the Ritz
    // vectors that have converged. 1-3 keep every phi_n (VB).
    const int rmode = o.reorth;
    const bool reorth = rmode != 0;
    const bool selective = rmode == 3;
    // the selective scheme looks for converged Ritz pairs after every step
    const int check = o.fixed_iter ? M : selective ? 1 : (o.check_every > 0 ? o.check_every : 10);
    constexpr int kMaxGood = 8;  // Ritz vectors kept by the selective scheme (full reorthogonalisation beyond)
    const double tol = o.tol > 0 ? o.tol : 1e-13;
    QMB_HIP(hipSetDevice(ctx->device));
    const unsigned rblocks = grid_for(n, kThreads * 4);
    {
        // the plan's vectors, refused up front when they cannot fit in device memory:
        // w, H phi, zero, then phi_{n-1}, phi_n or every kept phi_n with its coefficients
        const double need =
            (double)cnt * sizeof(double) * (3.0 + (reorth ? (double)M : 2.0) + (selective ? kMaxGood : 0.0)) +
            (reorth ? (double)M * rblocks * 2 * sizeof(double) : 0.0);
        size_t free_b = 0, total_b = 0;
        QMB_HIP(hipMemGetInfo(&free_b, &total_b));
        const double avail = (double)free_b + (double)ctx->pool.cached_bytes();
        if (need > avail) {
            char msg[256];
            std::snprintf(msg, sizeof msg,
                          "qmb_lanczos: the plan needs %.1f GB of device memory (%d kept Lanczos vectors of %.3f GB); "
                          "%.1f GB are free",
                          need / 1e9, reorth ? M : 2, (double)cnt * sizeof(double) / 1e9, avail / 1e9);
            throw Error{QMB_ERR_OOM, msg};
        }
    }
    LzWork wk;
    wk.init(ctx, M);
    DevBuf P(cnt * sizeof(double)), Y(cnt * sizeof(double)), Z(cnt * sizeof(double));
    DevBuf V0, V1, VB;
    if (reorth) {
        VB.alloc((size_t)M * cnt * sizeof(double));
    } else {
        V0.alloc(cnt * sizeof(double));
        V1.alloc(cnt * sizeof(double));
    }
    QMB_HIP(hipMemsetAsync(Z.p, 0, cnt * sizeof(double), ctx->stream));
    DevBuf rpart, rcoef;
    if (reorth) {
        rpart.alloc((size_t)M * rblocks * 2 * sizeof(double));
        rcoef.alloc((size_t)M * 2 * sizeof(double));
    }
    auto vcur = [&](int it) -> double * {
        if (reorth) return VB.as<double>() + (size_t)it * cnt;
        return (it & 1) ? V1.as<double>() : V0.as<double>();
    };
    auto vprev = [&](int it) -> const double * {
        if (it == 0) return Z.as<double>();
        return vcur(it - 1);
    };
    // start vector into P, 1/||P|| into invb[0]; beta[0] stays 0 (T's b[0])
    auto init_start = [&]() {
        if (o.start == QMB_START_USER) {
            QMB_REQUIRE(o.start_vec != nullptr, "qmb_lanczos: start == USER needs start_vec");
            QMB_HIP(hipMemcpyAsync(P.p, o.start_vec, cnt * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        } else if (o.start == QMB_START_DEVICE) {
            QMB_REQUIRE(o.start_vec != nullptr, "qmb_lanczos: start == DEVICE needs start_vec");
            QMB_HIP(hipMemcpyAsync(P.p, o.start_vec, cnt * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            hipLaunchKernelGGL(lz_random, dim3(grid_for(cnt, kThreads * 4)), dim3(kThreads), 0, ctx->stream, cnt,
                               o.seed, P.as<double>());
            QMB_HIP(hipGetLastError());
        }
        launch_norm(ctx, cnt, P.as<double>(), nullptr, 0.0, wk, wk.misc(), wk.invb());
    };
    const unsigned ublocks = grid_for(cnt, kThreads * 4);
    // Plain recurrence: the a[n] and b[n+1] one-block finishes are folded into
    // the next kernel (every block re-sums the fixed partial slots in the same
    // order, so the values are the same bit for bit). b_ready[n]: b[n], 1/b[n]
    // were written by an explicit finish (start vector, last step before a
    // convergence check); otherwise K1(n) derives them from K2(n-1)'s partials.
    const bool fuse = rmode == 0;
    double *part1 = wk.partials.as<double>(), *part2 = wk.partials.as<double>() + kMaxBlocks;
    const unsigned sblocks = spmv_blocks(e);
    std::vector<char> b_ready(M + 1, 0);
    b_ready[0] = 1;
    auto k1 = [&](int it, hipEvent_t after) {
        if (fuse) {
            FusedBeta fb;
            fb.partials = part1;
            if (!b_ready[it]) {
                fb.bpart = part2;
                fb.n_bpart = (int)ublocks;
                fb.beta_out = wk.beta() + it;
                fb.invb_out = wk.invb() + it;
            }
            launch_spmv(ctx, e, P.as<double>(), Y.as<double>(), true, wk.invb() + it, vcur(it), &wk,
                        wk.alpha() + it, after, &fb);
            return;
        }
        launch_spmv(ctx, e, P.as<double>(), Y.as<double>(), true, wk.invb() + it, vcur(it), &wk, wk.alpha() + it,
                    after);
    };
    // Partial reorthogonalisation (H. D. Simon, Math. Comp. 42 (1984) 115-142):
    // om_cur[k] estimates <phi_j, phi_k> through the recurrence the Lanczos
    // vectors themselves obey,
    //   b[j+1] om'[k] = b[k+1] om[k+1] + (a[k] - a[j]) om[k] + b[k] om[k-1] - b[j] om_prev[k],
    // plus a rounding term of size eps1 = eps sqrt(n) per step. When an estimate
    // passes sqrt(eps), the new vector and the one after it are orthogonalised
    // against every kept phi and the estimates restart at eps1.
    const double eps1 = 2.220446049250313e-16 * std::sqrt((double)n);
    std::vector<double> om_prev, om_cur, om_next, pa, pb;
    if (rmode == 2) {
        om_prev.assign(M + 2, 0.0);
        om_cur.assign(M + 2, 0.0);
        om_next.assign(M + 2, 0.0);
        om_cur[0] = 1.0;
        pa.assign(M + 1, 0.0);
        pb.assign(M + 2, 0.0);
    }
    bool pro_second = false;
    int n_reorth = 0;
    auto pro_decide = [&](int it) -> bool {
        // a[it], b[it+1] of this step (one small read per step)
        double *st = ctx->host_stage<double>(2 * sizeof(double));
        QMB_HIP(hipMemcpyAsync(st, wk.alpha() + it, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipMemcpyAsync(st + 1, wk.beta() + it + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        pa[it] = st[0];
        pb[it + 1] = st[1];
        const int j = it;
        const double bj1 = pb[j + 1];
        if (!(bj1 > 0.0)) return false;  // breakdown: nothing to orthogonalise
        double worst = 0.0;
        for (int k = 0; k < j; ++k) {
            double t = pb[k + 1] * om_cur[k + 1] + (pa[k] - pa[j]) * om_cur[k] - pb[j] * om_prev[k];
            if (k > 0) t += pb[k] * om_cur[k - 1];
            t /= bj1;
            t += (t >= 0.0 ? eps1 : -eps1);
            om_next[k] = t;
            worst = std::max(worst, std::fabs(t));
        }
        om_next[j] = eps1;
        om_next[j + 1] = 1.0;
        const bool run = pro_second || worst > 1.4901161193847656e-08;  // sqrt(eps)
        pro_second = run && !pro_second;
        if (run)
            for (int k = 0; k <= j; ++k) om_next[k] = eps1;
        std::swap(om_prev, om_cur);
        std::swap(om_cur, om_next);
        return run;
    };
    // w (in P) -= sum_j <V_j, w> V_j over nv vectors of stride cnt, classical
    // Gram-Schmidt; pass 1 only where pass 0 cancelled most of the norm (the flag
    // in misc[3] is set by pass 0's finish on the device); the new norm goes to
    // b[bslot], 1/b[bslot]
    auto gram_schmidt = [&](const double *V, int nv, int bslot) {
        double *gate = wk.misc() + 3;
        for (int pass = 0; pass < 2; ++pass) {
            const double *g = pass ? gate : nullptr;
            auto kd = cplx ? reorth_dots<true> : reorth_dots<false>;
            hipLaunchKernelGGL(kd, dim3(rblocks, (nv + kDotGroup - 1) / kDotGroup), dim3(kThreads), 0, ctx->stream,
                               n, (int64_t)cnt, nv, V, P.as<double>(), rpart.as<double>(), (int)rblocks, g);
            QMB_HIP(hipGetLastError());
            hipLaunchKernelGGL(reorth_reduce, dim3(nv), dim3(kThreads), 0, ctx->stream, rpart.as<double>(),
                               (int)rblocks, rcoef.as<double>(), g);
            QMB_HIP(hipGetLastError());
            auto ka = cplx ? reorth_apply<true> : reorth_apply<false>;
            hipLaunchKernelGGL(ka, dim3(rblocks), dim3(kThreads), 0, ctx->stream, n, (int64_t)cnt, nv, V,
                               rcoef.as<double>(), P.as<double>(), wk.partials.as<double>(), g);
            QMB_HIP(hipGetLastError());
            launch_finish(ctx, wk.partials.as<double>(), rblocks, wk.beta() + bslot, wk.invb() + bslot, pass ? 1 : 2,
                          gate);
        }
    };
    // Selective orthogonalisation (B. N. Parlett and D. S. Scott, Math. Comp. 33
    // (1979) 217-238): orthogonality is lost only towards converged Ritz vectors
    // (C. C. Paige), so once a Ritz pair of T is good, i.e. its residual
    // b[m] |s_m| is at most sqrt(eps) ||T||, its Ritz vector y = sum_k s_k phi_k
    // is formed once from the kept vectors and every later Lanczos vector is
    // orthogonalised against the good vectors. The lowest max(4, n_eigs + 1) and
    // the four highest Ritz pairs are watched (the extremes converge first);
    // beyond kMaxGood good vectors every vector is orthogonalised against all
    // kept ones.
    DevBuf GV, dzs;
    std::vector<double> good_theta, zs(selective ? (size_t)kMaxGood * M : 0);
    bool so_full = false;
    // finish: write b[it+1], 1/b[it+1] now (needed by a convergence check)
    auto k2 = [&](int it, bool finish) {
        if (fuse) {
            hipLaunchKernelGGL(lz_update, dim3(ublocks), dim3(kThreads), 0, ctx->stream, cnt, Y.as<double>(),
                               vcur(it), vprev(it), P.as<double>(), wk.alpha() + it, wk.beta() + it, part2, part1,
                               (int)sblocks);
            QMB_HIP(hipGetLastError());
            if (finish) {
                launch_finish(ctx, part2, ublocks, wk.beta() + it + 1, wk.invb() + it + 1, 1);
                b_ready[it + 1] = 1;
            }
            return;
        }
        hipLaunchKernelGGL(lz_update, dim3(ublocks), dim3(kThreads), 0, ctx->stream, cnt, Y.as<double>(), vcur(it),
                           vprev(it), P.as<double>(), wk.alpha() + it, wk.beta() + it, wk.partials.as<double>(),
                           nullptr, 0);
        QMB_HIP(hipGetLastError());
        launch_finish(ctx, wk.partials.as<double>(), ublocks, wk.beta() + it + 1, wk.invb() + it + 1, 1);
        const bool run = rmode == 1 || (rmode == 2 && pro_decide(it)) ||
                         (selective && (so_full || !good_theta.empty()));
        if (run) {
            ++n_reorth;
            if (selective && !so_full)
                gram_schmidt(GV.as<double>(), (int)good_theta.size(), it + 1);
            else
                gram_schmidt(VB.as<double>(), it + 1, it + 1);
            if (rmode == 2) {
                // the orthogonalised vector has a new norm b[it+1]
                double *st = ctx->host_stage<double>(sizeof(double));
                QMB_HIP(hipMemcpyAsync(st, wk.beta() + it + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
                QMB_HIP(hipStreamSynchronize(ctx->stream));
                pb[it + 1] = st[0];
            }
        }
    };

    // profiling events: one pair per SpMV, two check windows' worth (the first
    // step of a window may be queued while the previous window is evaluated)
    std::vector<hipEvent_t> ev;
    struct EventVec {
        std::vector<hipEvent_t> &v;
        ~EventVec() {
            for (auto e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } ev_guard{ev};
    if (o.profile) {
        ev.resize(4 * (size_t)check, nullptr);
        for (auto &x : ev) QMB_HIP(hipEventCreate(&x));
    }
    struct EventHolder {
        hipEvent_t e = nullptr;
        ~EventHolder() {
            if (e) (void)hipEventDestroy(e);
        }
    } copy_done;
    QMB_HIP(hipEventCreateWithFlags(&copy_done.e, hipEventDisableTiming));
    const hipEvent_t ev_copy = copy_done.e;
    double ms_spmv = 0.0, ms_host = 0.0;
    int n_spmv = 0, n_timed = 0;

    std::vector<double> ha(M), hb(M + 1);
    std::vector<double> w(k_eigs + 1), zl(k_eigs + 1), z0(M);
    int it = 0;
    int m_eff = 0;
    bool converged = false, breakdown = false, stopped = false;
    // Convergence is tested after `next` steps: check_every at first, then the
    // number of steps the residual's observed geometric decay predicts it needs
    // to reach the tolerance (at least 1, at most check_every).
    int next = check, it_prev = -1;
    double q_prev = 0.0;
    auto fill_header = [&](LanczosFileHeader &h) {
        std::memset(&h, 0, sizeof h);
        std::memcpy(h.magic, kLanczosMagic, 8);
        h.version = 1;
        h.header_bytes = sizeof h;
        h.n_rows = n;
        h.nnz = e->nnz;
        h.value_type = e->value_type;
        h.reorth = rmode;
        h.max_iter = M;
        h.n_eigs = k_eigs;
        h.start = o.start;
        h.fixed_iter = o.fixed_iter;
        h.check_every = check;
        h.seed = o.seed;
        h.tol = tol;
    };
    auto write_checkpoint = [&]() {
        LanczosFileHeader h;
        fill_header(h);
        h.it = it;
        h.n_reorth = n_reorth;
        h.pro_second = pro_second ? 1 : 0;
        h.next = next;
        h.it_prev = it_prev;
        h.n_spmv = n_spmv;
        h.q_prev = q_prev;
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        const std::string tmp = std::string(ck->path) + ".tmp";
        {
            File out(tmp.c_str(), "wb");
            out.write(&h, sizeof h);
            write_device(out, wk.alpha(), (size_t)it * sizeof(double));
            write_device(out, wk.beta(), (size_t)(it + 1) * sizeof(double));
            write_device(out, wk.invb(), (size_t)(it + 1) * sizeof(double));
            write_device(out, P.p, cnt * sizeof(double));
            if (reorth)
                write_device(out, VB.p, (size_t)it * cnt * sizeof(double));
            else if (it > 0)
                write_device(out, vcur(it - 1), cnt * sizeof(double));
            if (rmode == 2) {
                out.write(om_prev.data(), om_prev.size() * sizeof(double));
                out.write(om_cur.data(), om_cur.size() * sizeof(double));
                out.write(pa.data(), pa.size() * sizeof(double));
                out.write(pb.data(), pb.size() * sizeof(double));
            }
            out.close();
        }
        // replace the previous checkpoint only once the new one is complete
        QMB_REQUIRE(std::rename(tmp.c_str(), ck->path) == 0, std::string("cannot rename ") + tmp);
    };
    if (ck && ck->resume) {
        File in(ck->path, "rb");
        LanczosFileHeader h, want;
        in.read(&h, sizeof h);
        fill_header(want);
        QMB_REQUIRE(std::memcmp(h.magic, kLanczosMagic, 8) == 0 && h.version == 1 && h.header_bytes == sizeof h,
                    "qmb_lanczos_checkpointed: not a Lanczos checkpoint");
        QMB_REQUIRE(h.n_rows == want.n_rows && h.nnz == want.nnz && h.value_type == want.value_type &&
                        h.reorth == want.reorth && h.max_iter == want.max_iter && h.n_eigs == want.n_eigs &&
                        h.start == want.start && h.fixed_iter == want.fixed_iter &&
                        h.check_every == want.check_every && h.seed == want.seed && h.tol == want.tol,
                    "qmb_lanczos_checkpointed: the checkpoint belongs to another matrix or other options");
        QMB_REQUIRE(h.it >= 1 && h.it < M && h.next >= 1 && h.next <= check,
                    "qmb_lanczos_checkpointed: corrupt checkpoint");
        it = h.it;
        n_reorth = h.n_reorth;
        pro_second = h.pro_second != 0;
        next = h.next;
        it_prev = h.it_prev;
        n_spmv = h.n_spmv;
        q_prev = h.q_prev;
        read_device(in, wk.alpha(), (size_t)it * sizeof(double));
        read_device(in, wk.beta(), (size_t)(it + 1) * sizeof(double));
        read_device(in, wk.invb(), (size_t)(it + 1) * sizeof(double));
        b_ready[it] = 1;  // b[it], 1/b[it] come from the file
        read_device(in, P.p, cnt * sizeof(double));
        if (reorth)
            read_device(in, VB.p, (size_t)it * cnt * sizeof(double));
        else
            read_device(in, vcur(it - 1), cnt * sizeof(double));
        if (rmode == 2) {
            in.read(om_prev.data(), om_prev.size() * sizeof(double));
            in.read(om_cur.data(), om_cur.size() * sizeof(double));
            in.read(pa.data(), pa.size() * sizeof(double));
            in.read(pb.data(), pb.size() * sizeof(double));
        }
    } else {
        init_start();
    }
    const int it_start = it;
    int last_ck = it;
    // Plain recurrence without checkpoints: while the host reads and solves the
    // tridiagonal matrix at a check, the next step is already queued (with an
    // explicit b finish, which gives the same values as the fused one), so the
    // device does not idle. Not after a check the decay prediction scheduled:
    // that one usually ends the run, and its queued step would be wasted.
    const bool can_queue_ahead = fuse && !o.fixed_iter && !(ck && ck->path);
    bool predicted = false;
    int it_q = it;  // steps queued so far
    int wpar = 0;   // event half of the current window
    // the staging buffer at its largest size up front: it is never reallocated
    // while a step is queued
    ctx->host_stage<double>((size_t)(2 * M + 1) * sizeof(double));
    // profile = k: steps 0, k, 2k, ... are timed
    auto timed = [&](int step) { return o.profile > 0 && step % o.profile == 0; };
    auto queue_step = [&](int step, int slot, int half, bool finish) {
        hipEvent_t *e2 = timed(step) ? &ev[(size_t)half * 2 * check + 2 * slot] : nullptr;
        if (e2) QMB_HIP(hipEventRecord(e2[0], ctx->stream));
        k1(step, e2 ? e2[1] : nullptr);
        k2(step, finish);
        it_q = step + 1;
    };
    while (it < M) {
        const int win0 = it;
        const int win1 = std::min(M, it + next);
        for (; it < win1; ++it)
            if (it >= it_q) queue_step(it, it - win0, wpar, it + 1 == win1);
        n_spmv += win1 - win0;
        // a[0..M) and b[0..it] are one span of the coefficient array: one copy
        // into pinned memory
        const size_t span = (size_t)M + it + 1;
        double *st = ctx->host_stage<double>(span * sizeof(double));
        QMB_HIP(hipMemcpyAsync(st, wk.alpha(), span * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipEventRecord(ev_copy, ctx->stream));
        if (can_queue_ahead && !predicted && it < M) queue_step(it, 0, wpar ^ 1, true);
        QMB_HIP(hipEventSynchronize(ev_copy));
        std::copy(st, st + it, ha.begin());
        std::copy(st + M, st + M + it + 1, hb.begin());
        if (o.profile) {
            for (int q = 0; q < it - win0; ++q) {
                if (!timed(win0 + q)) continue;
                float t = 0.f;
                const size_t i0 = (size_t)wpar * 2 * check + 2 * q;
                QMB_HIP(hipEventElapsedTime(&t, ev[i0], ev[i0 + 1]));
                ms_spmv += t;
                ++n_timed;
            }
        }
        wpar ^= 1;
        Timer th;
        // breakdown: an (almost) zero b[j] ends the Krylov space at j
        double tsc = 0.0;
        for (int j = 0; j < it; ++j) tsc = std::max(tsc, std::fabs(ha[j]) + hb[j] + hb[j + 1]);
        m_eff = it;
        for (int j = 1; j <= it; ++j) {
            if (!(hb[j] > 1e-13 * std::max(tsc, 1e-300))) {
                m_eff = j;
                breakdown = true;
                break;
            }
        }
        const int kk = std::min(k_eigs + 1, m_eff);
        tridiag_lowest(m_eff, ha.data(), hb.data(), kk, w.data(), zl.data(), z0.data());
        ms_host += th.ms();
        if (breakdown) {
            converged = m_eff >= k_eigs;
            break;
        }
        if (o.fixed_iter) break;
        if (m_eff >= k_eigs) {
            bool ok = true;
            double q = 0.0;  // worst residual / its target over the wanted eigenvalues
            const double sq = std::sqrt(tol);
            for (int j = 0; j < k_eigs; ++j) {
                const double sc = std::max(1.0, std::fabs(w[j]));
                const double r = std::fabs(hb[m_eff] * zl[j]);
                double gap = 1.0;
                if (j + 1 < kk) gap = w[j + 1] - w[j];
                if (j > 0) gap = std::min(gap, w[j] - w[j - 1]);
                ok = ok && r <= sq * sc && r * r / std::max(gap, sq) <= tol * sc;
                const double target = std::min(sq * sc, std::sqrt(tol * sc * std::max(gap, sq)));
                q = std::max(q, r / target);
            }
            if (ok) {
                converged = true;
                break;
            }
            next = check;
            if (it_prev >= 0 && q > 1.0 && q < q_prev) {
                const double rate = std::log(q_prev / q) / (double)(it - it_prev);  // decay per step
                const double steps = std::ceil(std::log(q) / rate);
                if (steps < (double)check) next = std::max(1, (int)steps);
            }
            predicted = next < check;
            q_prev = q;
            it_prev = it;
        }
        if (selective && it < M) {
            // good Ritz pairs of T_m among the lowest max(4, n_eigs + 1) and the
            // four highest
            Timer ts;
            const int m = m_eff;
            const int kl = std::min(std::max(4, k_eigs + 1), m), kh = std::min(4, m);
            std::vector<double> wl(kl), zlo(kl), wh(kh), zhi(kh);
            tridiag_lowest(m, ha.data(), hb.data(), kl, wl.data(), zlo.data(), nullptr);
            tridiag_highest(m, ha.data(), hb.data(), kh, wh.data(), zhi.data());
            const double tnorm = std::max(std::fabs(wl[0]), std::fabs(wh[0]));
            const double thr = 1.4901161193847656e-08 * tnorm;  // sqrt(eps) ||T||
            const bool was_full = so_full;
            int added = 0;
            auto consider = [&](double theta, double zlast) {
                if (so_full || !(std::fabs(hb[m] * zlast) <= thr)) return;
                for (const double g : good_theta)
                    if (std::fabs(theta - g) <= 1e-9 * tnorm) return;
                if ((int)good_theta.size() == kMaxGood) {
                    so_full = true;
                    return;
                }
                if (!GV.p) {
                    GV.alloc((size_t)kMaxGood * cnt * sizeof(double));
                    dzs.alloc((size_t)kMaxGood * M * sizeof(double));
                }
                const int slot = (int)good_theta.size();
                tridiag_vector(m, ha.data(), hb.data(), theta, zs.data() + (size_t)slot * M);
                double *dz = dzs.as<double>() + (size_t)slot * M;
                QMB_HIP(hipMemcpyAsync(dz, zs.data() + (size_t)slot * M, (size_t)m * sizeof(double),
                                       hipMemcpyHostToDevice, ctx->stream));
                hipLaunchKernelGGL(basis_combine, dim3(grid_for(cnt, kThreads * 4)), dim3(kThreads), 0, ctx->stream,
                                   cnt, (int64_t)cnt, m, VB.as<double>(), dz, GV.as<double>() + (size_t)slot * cnt);
                QMB_HIP(hipGetLastError());
                good_theta.push_back(theta);
                ++added;
            };
            for (int j = 0; j < kl; ++j) consider(wl[j], zlo[j]);
            for (int j = 0; j < kh; ++j) consider(wh[j], zhi[j]);
            // the current w becomes phi_m at the next step: orthogonalise it now
            // against what changed (every later step does so in k2)
            if ((so_full && !was_full) || (!so_full && added)) {
                ++n_reorth;
                if (so_full)
                    gram_schmidt(VB.as<double>(), m, m);
                else
                    gram_schmidt(GV.as<double>(), (int)good_theta.size(), m);
            }
            ms_host += ts.ms();
            // the host coefficients (zs) are read by the transfers until here
            QMB_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (ck && ck->path && it < M) {
            const bool stop = ck->stop_after > 0 && it - it_start >= ck->stop_after;
            if (stop || (ck->every > 0 && it - last_ck >= ck->every)) {
                write_checkpoint();
                last_ck = it;
            }
            if (stop) {
                stopped = true;
                break;
            }
        }
    }
    // a step queued ahead of the last check is not part of the result; let it
    // drain before the buffers and events go
    if (it_q > it) QMB_HIP(hipStreamSynchronize(ctx->stream));
    if (o.fixed_iter) converged = true;
    const int m = m_eff;
    const int kk = std::min(k_eigs, m);
    for (int j = 0; j < k_eigs; ++j) evals[j] = j < kk ? w[j] : std::nan("");
    if (krylov) {
        for (int j = 0; j < M; ++j) krylov[j] = j < m ? z0[j] : 0.0;
    }
    double eig_res = 0.0;
    if (o.want_eigvec && !stopped) {
        // rebuild x = sum_j z0[j] phi_j (the reference's second pass, lanczos.cu:31-104)
        DevBuf X(cnt * sizeof(double));
        if (reorth) {
            DevBuf dz(m * sizeof(double));
            QMB_HIP(hipMemcpyAsync(dz.p, z0.data(), m * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(basis_combine, dim3(grid_for(cnt, kThreads * 4)), dim3(kThreads), 0, ctx->stream, cnt,
                               (int64_t)cnt, m, VB.as<double>(), dz.as<double>(), X.as<double>());
            QMB_HIP(hipGetLastError());
        } else {
            QMB_HIP(hipMemsetAsync(X.p, 0, cnt * sizeof(double), ctx->stream));
            QMB_HIP(hipMemsetAsync(Z.p, 0, cnt * sizeof(double), ctx->stream));
            init_start();
            std::fill(b_ready.begin(), b_ready.end(), 0);
            b_ready[0] = 1;
            for (int j = 0; j < m; ++j) {
                k1(j, nullptr);
                hipLaunchKernelGGL(vec_axpy, dim3(grid_for(cnt, kThreads * 4)), dim3(kThreads), 0, ctx->stream, cnt,
                                   z0[j], vcur(j), X.as<double>());
                QMB_HIP(hipGetLastError());
                k2(j, false);
            }
        }
        launch_spmv(ctx, e, X.as<double>(), Y.as<double>(), false, nullptr, nullptr, nullptr, nullptr);
        launch_norm(ctx, cnt, Y.as<double>(), X.as<double>(), evals[0], wk, wk.misc() + 1, wk.misc() + 2);
        QMB_HIP(hipMemcpyAsync(&eig_res, wk.misc() + 1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        if (eigvec)
            QMB_HIP(hipMemcpyAsync(eigvec, X.p, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        n_spmv += m + 1;
    }
    if (rep) {
        std::memset(rep, 0, sizeof(*rep));
        rep->iters = m;
        rep->converged = converged ? 1 : 0;
        rep->breakdown = breakdown ? 1 : 0;
        rep->n_spmv = n_spmv;
        rep->n_reorth = n_reorth;
        for (int j = 0; j < std::min(k_eigs, 8); ++j)
            rep->residual[j] = j < kk ? std::fabs(hb[m] * zl[j]) : std::nan("");
        rep->eig_residual = eig_res;
        rep->ms_total = wall.ms();
        rep->ms_spmv = ms_spmv;
        rep->n_spmv_timed = n_timed;
        rep->ms_host = ms_host;
    }
    return converged ? QMB_OK : QMB_ERR_NOT_CONVERGED;
}

// Reference-layout host arrays -> SELL-64 device ELL.
qmb_ell *ell_from_reference(qmb_ctx *ctx, int64_t n, int W, const uint32_t *ind, const double *val,
                            const int16_t *track, bool cplx) {
    auto e = std::make_unique<qmb_ell>();
    e->ctx = ctx;
    e->value_type = cplx ? QMB_COMPLEX : QMB_REAL;
    e->max_terms = W;
    e->n_rows = n;
    e->n_chunks = (n + kChunk - 1) / kChunk;
    const int vs = cplx ? 2 : 1;
    std::vector<int64_t> cp(e->n_chunks + 1, 0);
    std::vector<uint16_t> rl(std::max<int64_t>(n, 1), 0);
    int64_t nnz = 0;
    int mx = 0;
    for (int64_t c = 0; c < e->n_chunks; ++c) {
        int L = 0;
        for (int64_t r = c * kChunk; r < std::min<int64_t>(n, (c + 1) * kChunk); ++r) {
            QMB_REQUIRE(track[r] >= 0 && track[r] <= W, "reference ELL: track out of range");
            rl[r] = (uint16_t)track[r];
            L = std::max(L, (int)track[r]);
            nnz += track[r];
        }
        mx = std::max(mx, L);
        cp[c + 1] = cp[c] + (int64_t)kChunk * L;
    }
    e->stored = cp[e->n_chunks];
    e->nnz = nnz;
    e->max_row_len = mx;
    std::vector<int32_t> cl(std::max<int64_t>(e->stored, 1), 0);
    std::vector<double> vl((size_t)std::max<int64_t>(e->stored, 1) * vs, 0.0);
    for (int64_t r = 0; r < n; ++r) {
        const int64_t c = r / kChunk, lane = r % kChunk;
        const int L = (int)((cp[c + 1] - cp[c]) / kChunk);
        for (int s = 0; s < L; ++s) {
            const int64_t pos = cp[c] + (int64_t)s * kChunk + lane;
            if (s < track[r]) {
                const uint32_t j = ind[(size_t)r * W + s];
                QMB_REQUIRE((int64_t)j < n, "reference ELL: column out of range");
                cl[pos] = (int32_t)j;
                for (int q = 0; q < vs; ++q) vl[(size_t)pos * vs + q] = val[((size_t)r * W + s) * vs + q];
            } else {
                cl[pos] = (int32_t)r;
            }
        }
    }
    e->chunk_ptr.alloc(cp.size() * sizeof(int64_t));
    e->row_len.alloc(rl.size() * sizeof(uint16_t));
    e->col.alloc(cl.size() * sizeof(int32_t));
    e->val.alloc(vl.size() * sizeof(double));
    QMB_HIP(hipMemcpy(e->chunk_ptr.p, cp.data(), cp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    QMB_HIP(hipMemcpy(e->row_len.p, rl.data(), rl.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    QMB_HIP(hipMemcpy(e->col.p, cl.data(), cl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    QMB_HIP(hipMemcpy(e->val.p, vl.data(), vl.size() * sizeof(double), hipMemcpyHostToDevice));
    return e.release();
}

// Reference-layout DEVICE arrays -> SELL-64 device ELL, on the device: row
// lengths from track (checked to lie in [0, W]), chunk offsets by the device
// scan, then one thread per row copies its slots (checked columns) into the
// slot-major chunks. The caller's arrays are only read.
__global__ __launch_bounds__(256) void ref_row_len(const int16_t *__restrict__ track, int64_t n, int W,
                                                   uint16_t *__restrict__ row_len, int *__restrict__ err) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int t = track[r];
        const bool ok = t >= 0 && t <= W;
        row_len[r] = ok ? (uint16_t)t : (uint16_t)0;
        if (!ok) *err = 1;
    }
}

__global__ __launch_bounds__(256) void ref_fill(const uint32_t *__restrict__ ind, const double *__restrict__ val,
                                                int64_t n, int W, int vs, const int64_t *__restrict__ chunk_ptr,
                                                const uint16_t *__restrict__ row_len, int32_t *__restrict__ col,
                                                double *__restrict__ vout, int *__restrict__ err) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r / kChunk, lane = r % kChunk;
        const int64_t base = chunk_ptr[c];
        const int L = (int)((chunk_ptr[c + 1] - base) / kChunk);
        const int len = row_len[r];
        for (int s = 0; s < L; ++s) {
            const int64_t pos = base + (int64_t)s * kChunk + lane;
            int32_t j = (int32_t)r;
            if (s < len) {
                const uint32_t jj = ind[(size_t)r * W + s];
                if ((int64_t)jj < n)
                    j = (int32_t)jj;
                else
                    *err = 2;
            }
            col[pos] = j;
            for (int q = 0; q < vs; ++q) vout[(size_t)pos * vs + q] = s < len ? val[((size_t)r * W + s) * vs + q] : 0.0;
        }
    }
}

qmb_ell *ell_from_reference_device(qmb_ctx *ctx, int64_t n, int W, const uint32_t *ind, const double *val,
                                   const int16_t *track, bool cplx) {
    auto e = std::make_unique<qmb_ell>();
    e->ctx = ctx;
    e->value_type = cplx ? QMB_COMPLEX : QMB_REAL;
    e->max_terms = W;
    e->n_rows = n;
    e->n_chunks = (n + kChunk - 1) / kChunk;
    const int vs = cplx ? 2 : 1;
    e->row_len.alloc((size_t)n * sizeof(uint16_t));
    e->chunk_ptr.alloc((size_t)(e->n_chunks + 1) * sizeof(int64_t));
    DevBuf err(sizeof(int));
    const int zero = 0;
    QMB_HIP(hipMemcpyAsync(err.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    const int blocks = (int)std::min<int64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(ref_row_len, dim3(blocks), dim3(256), 0, ctx->stream, track, n, W, e->row_len.as<uint16_t>(),
                       err.as<int>());
    QMB_HIP(hipGetLastError());
    int herr = 0, mx = 0;
    int64_t nnz = 0;
    e->stored = scan_chunks_i64(ctx, e->row_len.as<uint16_t>(), n, e->chunk_ptr.as<int64_t>(), e->n_chunks, &nnz,
                                &mx, err.as<int>(), &herr);
    QMB_REQUIRE(herr == 0, "reference ELL: track out of range [0, max_terms]");
    e->nnz = nnz;
    e->max_row_len = mx;
    e->col.alloc((size_t)std::max<int64_t>(e->stored, 1) * sizeof(int32_t));
    e->val.alloc((size_t)std::max<int64_t>(e->stored, 1) * vs * sizeof(double));
    hipLaunchKernelGGL(ref_fill, dim3(blocks), dim3(256), 0, ctx->stream, ind, val, n, W, vs,
                       e->chunk_ptr.as<int64_t>(), e->row_len.as<uint16_t>(), e->col.as<int32_t>(),
                       e->val.as<double>(), err.as<int>());
    QMB_HIP(hipGetLastError());
    QMB_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    QMB_HIP(hipStreamSynchronize(ctx->stream));
    QMB_REQUIRE(herr == 0, "reference ELL: column out of range");
    return e.release();
}

// Fixed NIter steps from the seeded start on a reference-layout ELL: the
// reference free function's semantics (lanczos.cu:128-317).
double eigvals_ref_run(qmb_ctx *ctx, qmb_ell *e, double *E_tmp, int NIter, int nobv, long long seed,
                       double *eig_vec_Krylov_out, int exportKrylovVec, int checkEigState) {
    qmb_lanczos_opts o;
    qmb_lanczos_default_opts(&o);
    o.max_iter = NIter;
    o.fixed_iter = 1;
    o.seed = (uint64_t)seed;
    o.want_eigvec = checkEigState ? 1 : 0;
    std::vector<double> kr(NIter);
    qmb_lanczos_report rep;
    double ev = 0.0;
    qmb_status r = run_lanczos(ctx, e, o, &ev, kr.data(), nullptr, &rep);
    if (r != QMB_OK && r != QMB_ERR_NOT_CONVERGED) throw Error{r, qmb_last_error()};
    *E_tmp = ev;
    if (exportKrylovVec && eig_vec_Krylov_out)
        for (int i = 0; i < NIter; ++i) eig_vec_Krylov_out[i] = kr[i];
    return checkEigState ? rep.eig_residual / (double)nobv : 0.0;
}

qmb_ctx *default_ctx() {
    static std::once_flag once;
    static qmb_ctx *ctx = nullptr;
    static qmb_status st = QMB_OK;
    std::call_once(once, [] { st = qmb_ctx_create(0, &ctx); });
    if (st != QMB_OK) throw Error{st, std::string("default context: ") + qmb_last_error()};
    return ctx;
}

void check_opts(const qmb_lanczos_opts &o) {
    QMB_REQUIRE(o.max_iter >= 1 && o.max_iter <= 100000, "qmb_lanczos: max_iter must be in [1, 100000]");
    QMB_REQUIRE(o.n_eigs >= 1 && o.n_eigs <= 8, "qmb_lanczos: n_eigs must be in [1, 8]");
    QMB_REQUIRE(o.reorth >= 0 && o.reorth <= 3, "qmb_lanczos: reorth must be 0, 1, 2 or 3");
    QMB_REQUIRE(o.n_eigs == 1 || o.reorth, "qmb_lanczos: n_eigs > 1 needs reorth = 1, 2 or 3");
    QMB_REQUIRE(o.reorth != 3 || !o.fixed_iter,
                "qmb_lanczos: the selective scheme (reorth = 3) tests every step, not with fixed_iter");
    QMB_REQUIRE(o.start == QMB_START_RANDOM || o.start == QMB_START_USER || o.start == QMB_START_DEVICE,
                "qmb_lanczos: bad start mode");
}

}  // namespace

extern "C" {

void qmb_lanczos_default_opts(qmb_lanczos_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_iter = 300;
    o->n_eigs = 1;
    o->check_every = 10;
    o->tol = 1e-13;
    o->seed = 2212456ULL;
}

qmb_status qmb_spmv_device(qmb_ctx *ctx, const qmb_ell *e, const double *x_dev, double *y_dev) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && e && x_dev && y_dev, "qmb_spmv_device: NULL argument");
        QMB_HIP(hipSetDevice(ctx->device));
        if (e->n_rows == 0) return;
        launch_spmv(ctx, e, x_dev, y_dev, false, nullptr, nullptr, nullptr, nullptr);
    });
}

qmb_status qmb_spmv_host(qmb_ctx *ctx, const qmb_ell *e, const double *x, double *y) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && e && x && y, "qmb_spmv_host: NULL argument");
        QMB_HIP(hipSetDevice(ctx->device));
        const int64_t cnt = e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1);
        if (cnt == 0) return;
        DevBuf dx(cnt * sizeof(double)), dy(cnt * sizeof(double));
        QMB_HIP(hipMemcpyAsync(dx.p, x, cnt * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        launch_spmv(ctx, e, dx.as<double>(), dy.as<double>(), false, nullptr, nullptr, nullptr, nullptr);
        QMB_HIP(hipMemcpyAsync(y, dy.p, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qmb_status qmb_spmv_benchmark(qmb_ctx *ctx, const qmb_ell *e, int reps, double *ms_per_spmv) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && e && ms_per_spmv && reps >= 1, "qmb_spmv_benchmark: bad argument");
        QMB_HIP(hipSetDevice(ctx->device));
        const int64_t cnt = std::max<int64_t>(1, e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1));
        DevBuf dx(cnt * sizeof(double)), dy(cnt * sizeof(double));
        hipLaunchKernelGGL(lz_random, dim3(grid_for(cnt, kThreads * 4)), dim3(kThreads), 0, ctx->stream, cnt,
                           (uint64_t)7, dx.as<double>());
        QMB_HIP(hipGetLastError());
        launch_spmv(ctx, e, dx.as<double>(), dy.as<double>(), false, nullptr, nullptr, nullptr, nullptr);
        std::vector<hipEvent_t> ev(2, nullptr);
        struct EventVec {
            std::vector<hipEvent_t> &v;
            ~EventVec() {
                for (auto x : v)
                    if (x) (void)hipEventDestroy(x);
            }
        } ev_guard{ev};
        QMB_HIP(hipEventCreate(&ev[0]));
        QMB_HIP(hipEventCreate(&ev[1]));
        QMB_HIP(hipEventRecord(ev[0], ctx->stream));
        for (int i = 0; i < reps; ++i)
            launch_spmv(ctx, e, dx.as<double>(), dy.as<double>(), false, nullptr, nullptr, nullptr, nullptr);
        QMB_HIP(hipEventRecord(ev[1], ctx->stream));
        QMB_HIP(hipEventSynchronize(ev[1]));
        float t = 0.f;
        QMB_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
        *ms_per_spmv = (double)t / reps;
    });
}

qmb_status qmb_lanczos(qmb_ctx *ctx, const qmb_ell *e, const qmb_lanczos_opts *opts, double *evals, double *krylov,
                       double *eigvec, qmb_lanczos_report *report) {
    qmb_status result = QMB_OK;
    qmb_status st = guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && e && opts && evals, "qmb_lanczos: NULL argument");
        QMB_REQUIRE(e->ctx == ctx, "qmb_lanczos: ELL belongs to another context");
        QMB_REQUIRE(e->n_rows > 0, "qmb_lanczos: empty matrix");
        check_opts(*opts);
        qmb_lanczos_opts o = *opts;
        if (o.max_iter > e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1) + 1 && !o.fixed_iter)
            o.max_iter = (int)std::max<int64_t>(1, e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1) + 1);
        result = run_lanczos(ctx, e, o, evals, krylov, eigvec, report);
        if (result == QMB_ERR_NOT_CONVERGED) set_error("qmb_lanczos: tolerance not met within max_iter");
    });
    return st != QMB_OK ? st : result;
}

qmb_status qmb_lanczos_checkpointed(qmb_ctx *ctx, const qmb_ell *e, const qmb_lanczos_opts *opts, const char *path,
                                    int every, int resume, int stop_after, double *evals, double *krylov,
                                    double *eigvec, qmb_lanczos_report *report) {
    qmb_status result = QMB_OK;
    qmb_status st = guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && e && opts && evals && path, "qmb_lanczos_checkpointed: NULL argument");
        QMB_REQUIRE(e->ctx == ctx, "qmb_lanczos_checkpointed: ELL belongs to another context");
        QMB_REQUIRE(e->n_rows > 0, "qmb_lanczos_checkpointed: empty matrix");
        QMB_REQUIRE(every >= 0 && stop_after >= 0 && (resume == 0 || resume == 1),
                    "qmb_lanczos_checkpointed: bad checkpoint arguments");
        check_opts(*opts);
        QMB_REQUIRE(opts->reorth != 3,
                    "qmb_lanczos_checkpointed: the selective scheme (reorth = 3) keeps no checkpoint state");
        qmb_lanczos_opts o = *opts;
        if (o.max_iter > e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1) + 1 && !o.fixed_iter)
            o.max_iter = (int)std::max<int64_t>(1, e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1) + 1);
        Checkpoint ck;
        ck.path = path;
        ck.every = every;
        ck.resume = resume;
        ck.stop_after = stop_after;
        result = run_lanczos(ctx, e, o, evals, krylov, eigvec, report, &ck);
        if (result == QMB_ERR_NOT_CONVERGED)
            set_error("qmb_lanczos_checkpointed: not converged (stopped, or max_iter reached)");
    });
    return st != QMB_OK ? st : result;
}

double qmb_lanczos_eigvals_ref(double *E_tmp, int NStates, int NIter, int nobv, int max_terms, long long seed,
                               const uint32_t *index_mat, const double *vals_mat, const int16_t *track_vec,
                               double *eig_vec_Krylov_out, int exportKrylovVec, int checkEigState, int is_complex,
                               qmb_status *status) {
    double ret = 0.0;
    qmb_status st = guarded([&] {
        QMB_REQUIRE(E_tmp && index_mat && vals_mat && track_vec, "qmb_lanczos_eigvals_ref: NULL argument");
        QMB_REQUIRE(NStates == 1, "qmb_lanczos_eigvals_ref: NStates must be 1 (as in the reference)");
        QMB_REQUIRE(NIter >= 1 && nobv >= 1 && max_terms >= 1, "qmb_lanczos_eigvals_ref: bad sizes");
        qmb_ctx *ctx = default_ctx();
        QMB_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<qmb_ell> e(
            ell_from_reference(ctx, nobv, max_terms, index_mat, vals_mat, track_vec, is_complex != 0));
        ret = eigvals_ref_run(ctx, e.get(), E_tmp, NIter, nobv, seed, eig_vec_Krylov_out, exportKrylovVec,
                              checkEigState);
    });
    if (status) *status = st;
    return st == QMB_OK ? ret : std::nan("");
}

double qmb_lanczos_eigvals_ref_device(double *E_tmp, int NStates, int NIter, int nobv, int max_terms, long long seed,
                                      const uint32_t *index_mat, const double *vals_mat, const int16_t *track_vec,
                                      double *eig_vec_Krylov_out, int exportKrylovVec, int checkEigState,
                                      int is_complex, qmb_status *status) {
    double ret = 0.0;
    qmb_status st = guarded([&] {
        QMB_REQUIRE(E_tmp && index_mat && vals_mat && track_vec, "qmb_lanczos_eigvals_ref_device: NULL argument");
        QMB_REQUIRE(NStates == 1, "qmb_lanczos_eigvals_ref_device: NStates must be 1 (as in the reference)");
        QMB_REQUIRE(NIter >= 1 && nobv >= 1 && max_terms >= 1 && max_terms <= 65535,
                    "qmb_lanczos_eigvals_ref_device: bad sizes");
        qmb_ctx *ctx = default_ctx();
        QMB_HIP(hipSetDevice(ctx->device));
        const size_t nw = (size_t)nobv * max_terms;
        require_device_range(ctx, index_mat, nw * sizeof(uint32_t), "qmb_lanczos_eigvals_ref_device: index_mat");
        require_device_range(ctx, vals_mat, nw * (is_complex ? 2 : 1) * sizeof(double),
                             "qmb_lanczos_eigvals_ref_device: vals_mat");
        require_device_range(ctx, track_vec, (size_t)nobv * sizeof(int16_t), "qmb_lanczos_eigvals_ref_device: track_vec");
        std::unique_ptr<qmb_ell> e(
            ell_from_reference_device(ctx, nobv, max_terms, index_mat, vals_mat, track_vec, is_complex != 0));
        ret = eigvals_ref_run(ctx, e.get(), E_tmp, NIter, nobv, seed, eig_vec_Krylov_out, exportKrylovVec,
                              checkEigState);
    });
    if (status) *status = st;
    return st == QMB_OK ? ret : std::nan("");
}

}  // extern "C"
