// Batched Lanczos over many independent sectors (no reference counterpart; the
// reference solves irreps one after another,
// Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:127-229).
//
// All instances advance in lockstep with one iteration counter on the device.
// Every phase is ONE launch covering every instance: a block task table maps
// each block to (instance, chunk or element range); each instance owns a
// contiguous block range, so its dot products are reduced deterministically by
// a finishing kernel with one block per instance. Instances with real and complex matrices form two
// families with their own kernel instantiations. A chunk of iterations is
// captured once as a HIP graph and replayed; between replays the host solves
// the tridiagonal problems and clears the active flag of converged instances
// (their blocks return at once). With reorthogonalisation the Krylov vectors are
// kept and every new vector is orthogonalised against them by classical
// Gram-Schmidt, a second time only for the instances whose first pass removed
// most of the norm (decided on the device, so the graph stays fixed).
// SpMV tile shape (chunks per block) is picked per family by timing the
// candidates on the actual batch (autotuner with a per-process cache).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>

#include "internal.hpp"
#include "reduce.hpp"
#include "sell.hpp"

namespace qmb {

struct BInst {
    const int64_t *chunk_ptr;
    const uint16_t *row_len;
    const int32_t *col;
    const double *val;
    const double *diag, *vtab;   // packed values (packed = 1): diagonal, value table(s)
    int packed, col_bits, re_bits;
    int64_t n_rows, n_chunks;
    double *P, *Y, *VB;          // w, H phi, Krylov basis [M][n*vs] (two slots, it & 1, without reorthogonalisation)
    double *alpha, *beta, *invb;  // [M], [M+1], [M+1]
    double *coef;                 // [M*2] reorthogonalisation coefficients
    double *gate;                 // [1] second Gram-Schmidt pass wanted (1) or not (0)
    double *pro;                  // [3] this step orthogonalises (1/0), second-step flag, steps that did
    double *omega;                // [3][M+2] orthogonality estimates (partial reorthogonalisation)
    int spmv_first, spmv_nblk;
    int vec_first, vec_nblk;
    int vb_ring;                  // 1: VB holds only phi_it and phi_{it-1} (plain recurrence)
};

// Slot of phi_it in an instance's Krylov storage.
__device__ inline size_t vb_slot(const BInst &I, int it) { return (size_t)(I.vb_ring ? (it & 1) : it); }

struct BTask {
    int inst;
    int pad;
    int64_t lo, hi;
};

// vector-phase elements (doubles) per block: 4 per thread (72 N=24 sectors:
// 8 per thread 3 % slower, 16 per thread 17 % slower, 2 per thread equal)
constexpr int kVecPerBlock = kThreads * 4;

// NT: matrix stream past the caches (nontemporal) or plain loads; U: slots per
// group of the packed row products. Both are picked by the autotuner.
template <bool CPLX, bool NT, int U>
__global__ __launch_bounds__(kThreads) void bk_spmv(const BInst *__restrict__ insts, const BTask *__restrict__ tasks,
                                                    const int *__restrict__ active, const int *__restrict__ it_p,
                                                    double *__restrict__ part) {
    __shared__ double lds[kWaves];
    const BTask t = tasks[blockIdx.x];
    if (!active[t.inst]) return;
    const BInst I = insts[t.inst];
    const int it = *it_p;
    const double s = I.invb[it];
    const int vs = CPLX ? 2 : 1;
    double *vout = I.VB + vb_slot(I, it) * I.n_rows * vs;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double dot = 0.0;
    for (int64_t c = t.lo + wv; c < t.hi; c += kWaves) {
        const int64_t r = c * 64 + lane;
        const int64_t base = I.chunk_ptr[c];
        const int L = (int)((I.chunk_ptr[c + 1] - base) >> 6);
        const int len = r < I.n_rows ? (int)I.row_len[r] : 0;
        const int64_t p0 = base + lane;
        if (!CPLX) {
            const double xr = r < I.n_rows ? I.P[r] : 0.0;
            const double acc = (I.packed ? sell_row_packed<NT, U>(reinterpret_cast<const uint32_t *>(I.col), I.vtab,
                                                                I.P, p0, L, len, I.col_bits,
                                                                r < I.n_rows ? I.diag[r] : 0.0, xr)
                                         : sell_row_real<NT>(I.col, I.val, I.P, p0, L, len)) *
                               s;
            if (r < I.n_rows) {
                I.Y[r] = acc;
                const double vr = xr * s;
                vout[r] = vr;
                dot += vr * acc;
            }
        } else {
            const double2 xr = r < I.n_rows ? reinterpret_cast<const double2 *>(I.P)[r] : make_double2(0.0, 0.0);
            double2 a = I.packed
                            ? sell_row_cplx_packed<NT, U>(reinterpret_cast<const uint32_t *>(I.col), I.vtab,
                                                   I.vtab + kMaxValues, I.P, p0, L, len, I.col_bits, I.re_bits,
                                                   r < I.n_rows ? reinterpret_cast<const double2 *>(I.diag)[r]
                                                                : make_double2(0.0, 0.0),
                                                   xr)
                            : sell_row_cplx<NT>(I.col, I.val, I.P, p0, L, len);
            a.x *= s;
            a.y *= s;
            if (r < I.n_rows) {
                reinterpret_cast<double2 *>(I.Y)[r] = a;
                double2 vr = reinterpret_cast<const double2 *>(I.P)[r];
                vr.x = vr.x * s;
                vr.y = vr.y * s;
                reinterpret_cast<double2 *>(vout)[r] = vr;
                dot += vr.x * a.x + vr.y * a.y;
            }
        }
    }
    const double tb = block_sum(dot, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tb;
}

// Second Gram-Schmidt pass only when the first removed more than this share
// of the norm (Daniel, Gragg, Kaufman, Stewart 1976).
constexpr double kDgks = 0.7071067811865476;

// One block per instance of the family: fixed-order sum of the instance's block
// partials. mode 0: alpha[it] = s; 1: beta[it+1] = sqrt(s), invb[it+1];
// 2 (start vector): invb[0] = 1/sqrt(s), beta[0] = 0; 3 (after the first
// Gram-Schmidt pass): as 1, and gate = 1 if sqrt(s) < kDgks * the old
// beta[it+1]; 4 (after the second pass): as 1 where the gate is 1.
__global__ __launch_bounds__(kThreads) void bk_finish(const BInst *__restrict__ insts, const int *__restrict__ active,
                                                      const int *__restrict__ it_p, const double *__restrict__ part,
                                                      int first_inst, int mode) {
    __shared__ double lds[kWaves];
    const int inst = first_inst + blockIdx.x;
    if (active && !active[inst]) return;
    const BInst I = insts[inst];
    if ((mode == 3 || mode == 4) && I.pro[0] == 0.0) return;
    if (mode == 4 && *I.gate == 0.0) return;
    const int it = it_p ? *it_p : 0;
    const int first = mode == 0 ? I.spmv_first : I.vec_first;
    const int nblk = mode == 0 ? I.spmv_nblk : I.vec_nblk;
    const double s = sum_partials(part + first, nblk, lds);
    if (threadIdx.x == 0) {
        if (mode == 0) {
            I.alpha[it] = s;
        } else if (mode == 1 || mode == 3 || mode == 4) {
            const double nb = sqrt(s);
            if (mode == 3) *I.gate = nb < kDgks * I.beta[it + 1] ? 1.0 : 0.0;
            I.beta[it + 1] = nb;
            I.invb[it + 1] = nb > 0.0 ? 1.0 / nb : 0.0;
        } else {
            const double nr = sqrt(s);
            I.invb[0] = nr > 0.0 ? 1.0 / nr : 0.0;
            I.beta[0] = 0.0;
        }
    }
}

// w = (-a) phi_it + (y + (-b) phi_{it-1}); ||w|| -> beta[it+1]. Ranges are in doubles.
__global__ __launch_bounds__(kThreads) void bk_update(const BInst *__restrict__ insts, const BTask *__restrict__ tasks,
                                                      const int *__restrict__ active, const int *__restrict__ it_p,
                                                      double *__restrict__ part, int vs) {
    __shared__ double lds[kWaves];
    const BTask t = tasks[blockIdx.x];
    if (!active[t.inst]) return;
    const BInst I = insts[t.inst];
    const int it = *it_p;
    const int64_t cnt = I.n_rows * vs;
    const double a = I.alpha[it], b = I.beta[it];
    const double *v = I.VB + vb_slot(I, it) * cnt;
    const double *vp = it > 0 ? I.VB + vb_slot(I, it - 1) * cnt : nullptr;
    double acc = 0.0;
    for (int64_t i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
        const double tt = vp ? I.Y[i] + (-b) * vp[i] : I.Y[i] + (-b) * 0.0;
        const double wi = (-a) * v[i] + tt;
        I.P[i] = wi;
        acc += wi * wi;
    }
    const double tb = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tb;
}

// Reorthogonalisation launch shape: blockIdx.y = 0..kJY-1 splits the Krylov
// index, and a block takes kJB vectors per pass over its range (w read once,
// one combined reduction for the kJB dots).
constexpr int kJY = 8;
constexpr int kJB = 4;

// part2[(j*nbv + block)*2 + {0,1}] = <V_j, w> over the block's range, j <= it.
// Per dot: thread-strided sums, wave shuffles, then the waves in order.
template <bool CPLX>
__global__ __launch_bounds__(kThreads) void bk_rdots(const BInst *__restrict__ insts, const BTask *__restrict__ tasks,
                                                     const int *__restrict__ active, const int *__restrict__ it_p,
                                                     double *__restrict__ part2, int nbv, int gated) {
    __shared__ double lds[kWaves * kJB * 2];
    const BTask t = tasks[blockIdx.x];
    if (!active[t.inst]) return;
    const BInst I = insts[t.inst];
    if (I.pro[0] == 0.0 || (gated && *I.gate == 0.0)) return;
    const int it = *it_p;
    const int vs = CPLX ? 2 : 1;
    const int64_t cnt = I.n_rows * vs;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j0 = (int)blockIdx.y * kJB; j0 <= it; j0 += kJY * kJB) {
        const int nj = min(kJB, it + 1 - j0);
        double sr[kJB], si[kJB];
#pragma unroll
        for (int u = 0; u < kJB; ++u) sr[u] = si[u] = 0.0;
        if (CPLX) {
            for (int64_t i = t.lo / 2 + threadIdx.x; i < t.hi / 2; i += blockDim.x) {
                const double wr = I.P[2 * i], wi = I.P[2 * i + 1];
#pragma unroll
                for (int u = 0; u < kJB; ++u) {
                    if (u < nj) {
                        const double *vj = I.VB + (size_t)(j0 + u) * cnt;
                        const double vr = vj[2 * i], vi = vj[2 * i + 1];
                        sr[u] += vr * wr + vi * wi;
                        si[u] += vr * wi - vi * wr;
                    }
                }
            }
        } else {
            for (int64_t i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
                const double w = I.P[i];
#pragma unroll
                for (int u = 0; u < kJB; ++u)
                    if (u < nj) sr[u] += I.VB[(size_t)(j0 + u) * cnt + i] * w;
            }
        }
#pragma unroll
        for (int u = 0; u < kJB; ++u) {
            sr[u] = wave_sum(sr[u]);
            if (CPLX) si[u] = wave_sum(si[u]);
        }
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < kJB; ++u) {
                lds[(wv * kJB + u) * 2] = sr[u];
                lds[(wv * kJB + u) * 2 + 1] = CPLX ? si[u] : 0.0;
            }
        __syncthreads();
        if ((int)threadIdx.x < 2 * nj) {
            const int u = (int)threadIdx.x >> 1, c = (int)threadIdx.x & 1;
            double tot = 0.0;
            for (int w = 0; w < kWaves; ++w) tot += lds[(w * kJB + u) * 2 + c];
            part2[((size_t)(j0 + u) * nbv + blockIdx.x) * 2 + c] = tot;
        }
        __syncthreads();
    }
}

// coef_j = sum of the instance's block partials in a fixed order (lane-strided
// sums, then wave shuffles): one wavefront per Krylov index j, gridDim.y blocks
// of kWaves wavefronts per instance.
__global__ __launch_bounds__(kThreads) void bk_rreduce(const BInst *__restrict__ insts, const int *__restrict__ active,
                                                       const int *__restrict__ it_p, const double *__restrict__ part2,
                                                       int nbv, int first_inst, int gated) {
    const int inst = first_inst + blockIdx.x;
    if (!active[inst]) return;
    const BInst I = insts[inst];
    if (I.pro[0] == 0.0 || (gated && *I.gate == 0.0)) return;
    const int it = *it_p;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = (int)blockIdx.y * kWaves + wv; j <= it; j += (int)gridDim.y * kWaves) {
        const double *pj = part2 + ((size_t)j * nbv + I.vec_first) * 2;
        double sr = 0.0, si = 0.0;
        for (int b = lane; b < I.vec_nblk; b += 64) {
            sr += pj[2 * b];
            si += pj[2 * b + 1];
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) {
            I.coef[2 * j] = sr;
            I.coef[2 * j + 1] = si;
        }
    }
}

// w -= sum_j coef_j V_j (j in order; four vectors loaded per step), ||w||^2 partials.
template <bool CPLX>
__global__ __launch_bounds__(kThreads) void bk_rapply(const BInst *__restrict__ insts, const BTask *__restrict__ tasks,
                                                      const int *__restrict__ active, const int *__restrict__ it_p,
                                                      double *__restrict__ part, int gated) {
    __shared__ double lds[kWaves];
    const BTask t = tasks[blockIdx.x];
    if (!active[t.inst]) return;
    const BInst I = insts[t.inst];
    if (I.pro[0] == 0.0 || (gated && *I.gate == 0.0)) return;
    const int it = *it_p;
    const int vs = CPLX ? 2 : 1;
    const int64_t cnt = I.n_rows * vs;
    double acc = 0.0;
    if (CPLX) {
        for (int64_t i = t.lo / 2 + threadIdx.x; i < t.hi / 2; i += blockDim.x) {
            double wr = I.P[2 * i], wi = I.P[2 * i + 1];
            int j = 0;
            for (; j + 4 <= it + 1; j += 4) {
                double vr[4], vi[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    vr[u] = I.VB[(size_t)(j + u) * cnt + 2 * i];
                    vi[u] = I.VB[(size_t)(j + u) * cnt + 2 * i + 1];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double cr = I.coef[2 * (j + u)], ci = I.coef[2 * (j + u) + 1];
                    wr -= cr * vr[u] - ci * vi[u];
                    wi -= cr * vi[u] + ci * vr[u];
                }
            }
            for (; j <= it; ++j) {
                const double cr = I.coef[2 * j], ci = I.coef[2 * j + 1];
                const double vr = I.VB[(size_t)j * cnt + 2 * i], vi = I.VB[(size_t)j * cnt + 2 * i + 1];
                wr -= cr * vr - ci * vi;
                wi -= cr * vi + ci * vr;
            }
            I.P[2 * i] = wr;
            I.P[2 * i + 1] = wi;
            acc += wr * wr + wi * wi;
        }
    } else {
        for (int64_t i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
            double wr = I.P[i];
            int j = 0;
            for (; j + 4 <= it + 1; j += 4) {
                double v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = I.VB[(size_t)(j + u) * cnt + i];
#pragma unroll
                for (int u = 0; u < 4; ++u) wr -= I.coef[2 * (j + u)] * v[u];
            }
            for (; j <= it; ++j) wr -= I.coef[2 * j] * I.VB[(size_t)j * cnt + i];
            I.P[i] = wr;
            acc += wr * wr;
        }
    }
    const double tb = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tb;
}

// Partial reorthogonalisation on the device (H. D. Simon, Math. Comp. 42 (1984)
// 115-142), one wavefront per instance of the family: the estimates
// om_{j+1,k} of <phi_{j+1}, phi_k> follow the Lanczos recurrence,
//   b[j+1] om_{j+1,k} = b[k+1] om_{j,k+1} + (a[k] - a[j]) om_{j,k} + b[k] om_{j,k-1} - b[j] om_{j-1,k},
// plus a rounding term eps1 = eps sqrt(n); when one passes sqrt(eps), this
// step and the next orthogonalise (pro[0] = 1) and the estimates restart at
// eps1. Rows of omega rotate by step parity (j+1, j, j-1 mod 3).
__global__ __launch_bounds__(64) void bk_pro(const BInst *__restrict__ insts, const int *__restrict__ active,
                                             const int *__restrict__ it_p, int first_inst, int M) {
    const int inst = first_inst + blockIdx.x;
    if (!active[inst]) return;
    const BInst I = insts[inst];
    const int j = *it_p, lane = threadIdx.x;
    const int64_t stride = M + 2;
    double *om_next = I.omega + (size_t)((j + 1) % 3) * stride;
    const double *om_cur = I.omega + (size_t)(j % 3) * stride;
    const double *om_prev = I.omega + (size_t)((j + 2) % 3) * stride;
    const double eps1 = 2.220446049250313e-16 * sqrt((double)I.n_rows);
    const double bj1 = I.beta[j + 1];
    double worst = 0.0;
    for (int k = lane; k < j; k += 64) {
        double t = I.beta[k + 1] * om_cur[k + 1] + (I.alpha[k] - I.alpha[j]) * om_cur[k] - I.beta[j] * om_prev[k];
        if (k > 0) t += I.beta[k] * om_cur[k - 1];
        t = bj1 > 0.0 ? t / bj1 : 0.0;
        t += t >= 0.0 ? eps1 : -eps1;
        om_next[k] = t;
        worst = fmax(worst, fabs(t));
    }
    for (int off = 32; off > 0; off >>= 1) worst = fmax(worst, __shfl_down(worst, off, 64));
    worst = __shfl(worst, 0, 64);
    const bool second = I.pro[1] != 0.0;
    const bool run = bj1 > 0.0 && (second || worst > 1.4901161193847656e-08);  // sqrt(eps)
    if (run)
        for (int k = lane; k <= j; k += 64) om_next[k] = eps1;
    if (lane == 0) {
        if (!run) om_next[j] = eps1;
        om_next[j + 1] = 1.0;
        I.pro[0] = run ? 1.0 : 0.0;
        I.pro[1] = (run && !second) ? 1.0 : 0.0;
        if (run) I.pro[2] += 1.0;
    }
}

__global__ void bk_advance(int *it_p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *it_p = *it_p + 1;
}

// start vectors and their norms (outside the graph)
__global__ __launch_bounds__(kThreads) void bk_start(const BInst *__restrict__ insts, const BTask *__restrict__ tasks,
                                                     double *__restrict__ part, uint64_t seed, int vs) {
    __shared__ double lds[kWaves];
    const BTask t = tasks[blockIdx.x];
    const BInst I = insts[t.inst];
    double acc = 0.0;
    for (int64_t i = t.lo + threadIdx.x; i < t.hi; i += blockDim.x) {
        const double u = splitmix_uniform(seed, (uint64_t)i);
        I.P[i] = u;
        acc += u * u;
    }
    const double tb = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = tb;
    (void)vs;
}

struct Family {
    bool cplx = false;
    std::vector<int> members;  // instance indices (global numbering)
    int first_inst = 0;        // position of the family's first instance in the device BInst array
    std::vector<BTask> spmv_tasks, vec_tasks;
    DevBuf d_spmv_tasks, d_vec_tasks;
    int spmv_first_block = 0, vec_first_block = 0;
    int cpb = 4;      // chunks per block
    int variant = 0;  // index into spmv_variants()
};

using SpmvKernel = void (*)(const BInst *, const BTask *, const int *, const int *, double *);

// The SpMV shapes the autotuner chooses among, per family: (stream policy,
// slots per group). Entry 0 is the streamed 8-slot form.
struct SpmvVariant {
    bool nt;
    int slots;
    SpmvKernel real, cplx;
};
static const std::vector<SpmvVariant> &spmv_variants() {
    static const std::vector<SpmvVariant> v = {
        {true, 8, bk_spmv<false, true, 8>, bk_spmv<true, true, 8>},
        {true, 4, bk_spmv<false, true, 4>, bk_spmv<true, true, 4>},
        {false, 4, bk_spmv<false, false, 4>, bk_spmv<true, false, 4>},
        {false, 8, bk_spmv<false, false, 8>, bk_spmv<true, false, 8>},
    };
    return v;
}

struct TileChoice {
    int cpb, variant;
};

// (complex, log2 rows, instances) -> tile choice; shared by every context of the
// process, hence the lock
static std::map<std::tuple<int, int, int>, TileChoice> g_tile_cache;
static std::mutex g_tile_mutex;

static bool tile_cache_get(const std::tuple<int, int, int> &key, TileChoice &c) {
    std::lock_guard<std::mutex> lock(g_tile_mutex);
    auto it = g_tile_cache.find(key);
    if (it == g_tile_cache.end()) return false;
    c = it->second;
    return true;
}

static void tile_cache_put(const std::tuple<int, int, int> &key, TileChoice c) {
    std::lock_guard<std::mutex> lock(g_tile_mutex);
    g_tile_cache[key] = c;
}

}  // namespace qmb

using namespace qmb;

namespace qmb {

// x = sum_j z[j] V[j * cnt + i] over the instance's kept Lanczos vectors (the
// single solver's basis combination, same order of the sum)
__global__ __launch_bounds__(kThreads) void bk_combine(int64_t cnt, int m, const double *__restrict__ V,
                                                       const double *__restrict__ z, double *__restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s = s + z[j] * V[(size_t)j * cnt + i];
        x[i] = s;
    }
}

// vecs: NULL (eigenvalues only) or one host array per instance
static void run_batched(qmb_ctx *ctx, int n, const qmb_ell *const *ells, const qmb_lanczos_opts *opts,
                        double *evals, double *const *vecs, qmb_lanczos_report *reports) {
    {
        QMB_REQUIRE(ctx && ells && opts && evals && n >= 1, "qmb_lanczos_batched: bad argument");
        const qmb_lanczos_opts o = *opts;
        QMB_REQUIRE(o.max_iter >= 1 && o.max_iter <= 5000, "qmb_lanczos_batched: max_iter must be in [1, 5000]");
        QMB_REQUIRE(o.n_eigs >= 1 && o.n_eigs <= 8, "qmb_lanczos_batched: n_eigs must be in [1, 8]");
        QMB_REQUIRE(o.reorth >= 0 && o.reorth <= 2, "qmb_lanczos_batched: reorth must be 0, 1 or 2");
        QMB_REQUIRE(o.n_eigs == 1 || o.reorth, "qmb_lanczos_batched: n_eigs > 1 needs reorth = 1 or 2");
        QMB_REQUIRE(o.start == QMB_START_RANDOM, "qmb_lanczos_batched: only the seeded random start");
        for (int i = 0; i < n; ++i) {
            QMB_REQUIRE(ells[i] && ells[i]->ctx == ctx, "qmb_lanczos_batched: ELL missing or from another context");
            QMB_REQUIRE(ells[i]->n_rows > 0, "qmb_lanczos_batched: empty matrix in the batch");
        }
        auto wall0 = std::chrono::steady_clock::now();
        QMB_HIP(hipSetDevice(ctx->device));
        const int chunk = o.check_every > 0 ? o.check_every : 10;
        const int M = ((o.max_iter + chunk - 1) / chunk) * chunk;
        const double tol = o.tol > 0 ? o.tol : 1e-13;
        const bool reorth = o.reorth != 0;
        const bool partial = o.reorth == 2;  // device-side decisions (bk_pro)
        const int k_eigs = o.n_eigs;

        // families: real first, then complex
        Family fam[2];
        fam[0].cplx = false;
        fam[1].cplx = true;
        for (int i = 0; i < n; ++i) fam[ells[i]->value_type == QMB_COMPLEX ? 1 : 0].members.push_back(i);
        std::vector<int> order;  // device instance position -> global index
        for (auto &f : fam) {
            f.first_inst = (int)order.size();
            for (int i : f.members) order.push_back(i);
        }
        // per-instance device buffers
        std::vector<BInst> hin(n);
        std::vector<DevBuf> bufs;
        bufs.reserve((size_t)n * 6);
        DevBuf scal((size_t)n * (3 * M + 2) * sizeof(double));
        DevBuf coefs((size_t)n * M * 2 * sizeof(double));
        DevBuf gates((size_t)n * sizeof(double));
        QMB_HIP(hipMemsetAsync(scal.p, 0, scal.bytes, ctx->stream));
        QMB_HIP(hipMemsetAsync(gates.p, 0, gates.bytes, ctx->stream));
        // orthogonality estimates: row 0 = om_0 = e_0, rows 1, 2 zero; flags:
        // full reorthogonalisation runs every step (pro[0] = 1 throughout)
        const size_t om_stride = (size_t)M + 2;
        DevBuf omegas(partial ? (size_t)n * 3 * om_stride * sizeof(double) : 8), pros((size_t)n * 3 * sizeof(double));
        {
            std::vector<double> hp((size_t)n * 3, 0.0);
            for (int i = 0; i < n; ++i) hp[(size_t)i * 3] = partial ? 0.0 : 1.0;
            QMB_HIP(hipMemcpy(pros.p, hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice));
            if (partial) {
                std::vector<double> ho((size_t)n * 3 * om_stride, 0.0);
                for (int i = 0; i < n; ++i) ho[(size_t)i * 3 * om_stride] = 1.0;
                QMB_HIP(hipMemcpy(omegas.p, ho.data(), ho.size() * sizeof(double), hipMemcpyHostToDevice));
            }
        }
        for (int pos = 0; pos < n; ++pos) {
            const qmb_ell *e = ells[order[pos]];
            const int vs = e->value_type == QMB_COMPLEX ? 2 : 1;
            const int64_t cnt = e->n_rows * vs;
            BInst &I = hin[pos];
            I.chunk_ptr = e->chunk_ptr.as<int64_t>();
            I.row_len = e->row_len.as<uint16_t>();
            I.col = e->col.as<int32_t>();
            I.val = e->val.as<double>();
            I.diag = e->diag.as<double>();
            I.vtab = e->vtab.as<double>();
            I.packed = e->packed;
            I.re_bits = e->re_bits;
            I.col_bits = e->col_bits;
            I.n_rows = e->n_rows;
            I.n_chunks = e->n_chunks;
            bufs.emplace_back(cnt * sizeof(double));
            I.P = bufs.back().as<double>();
            bufs.emplace_back(cnt * sizeof(double));
            I.Y = bufs.back().as<double>();
            // every phi_n is kept only when it is reorthogonalised against
            I.vb_ring = reorth ? 0 : 1;
            bufs.emplace_back((size_t)(reorth ? M : 2) * cnt * sizeof(double));
            I.VB = bufs.back().as<double>();
            double *sc = scal.as<double>() + (size_t)pos * (3 * M + 2);
            I.alpha = sc;
            I.beta = sc + M;
            I.invb = sc + 2 * M + 1;
            I.coef = coefs.as<double>() + (size_t)pos * M * 2;
            I.gate = gates.as<double>() + pos;
            I.pro = pros.as<double>() + 3 * (size_t)pos;
            I.omega = partial ? omegas.as<double>() + (size_t)pos * 3 * om_stride : nullptr;
        }
        // vector task tables (ranges in doubles, both families in one table per family)
        int total_vec_blocks = 0;
        for (auto &f : fam) {
            f.vec_first_block = total_vec_blocks;
            for (size_t q = 0; q < f.members.size(); ++q) {
                const int pos = f.first_inst + (int)q;
                const qmb_ell *e = ells[order[pos]];
                const int64_t cnt = e->n_rows * (f.cplx ? 2 : 1);
                const int64_t per = f.cplx ? 2 * kVecPerBlock : kVecPerBlock;  // keep complex pairs together
                hin[pos].vec_first = total_vec_blocks;
                int nb = 0;
                for (int64_t lo = 0; lo < cnt; lo += per, ++nb)
                    f.vec_tasks.push_back(BTask{pos, 0, lo, std::min<int64_t>(cnt, lo + per)});
                hin[pos].vec_nblk = nb;
                total_vec_blocks += nb;
            }
        }
        const int nbv = total_vec_blocks;
        auto build_spmv_tasks = [&](Family &f, int cpb, int &first_block_io) {
            f.spmv_tasks.clear();
            f.spmv_first_block = first_block_io;
            for (size_t q = 0; q < f.members.size(); ++q) {
                const int pos = f.first_inst + (int)q;
                hin[pos].spmv_first = first_block_io + (int)f.spmv_tasks.size();
                int nb = 0;
                for (int64_t lo = 0; lo < hin[pos].n_chunks; lo += cpb, ++nb)
                    f.spmv_tasks.push_back(BTask{pos, 0, lo, std::min<int64_t>(hin[pos].n_chunks, lo + cpb)});
                hin[pos].spmv_nblk = nb;
            }
            first_block_io += (int)f.spmv_tasks.size();
        };
        // device tables
        DevBuf d_inst((size_t)n * sizeof(BInst)), d_active((size_t)n * sizeof(int)), d_it(sizeof(int));
        std::vector<int> active(n, 1);
        const int zero = 0;
        QMB_HIP(hipMemcpyAsync(d_it.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        QMB_HIP(hipMemcpyAsync(d_active.p, active.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        for (auto &f : fam) {
            if (f.vec_tasks.empty()) continue;
            f.d_vec_tasks.alloc(f.vec_tasks.size() * sizeof(BTask));
            QMB_HIP(hipMemcpy(f.d_vec_tasks.p, f.vec_tasks.data(), f.vec_tasks.size() * sizeof(BTask),
                              hipMemcpyHostToDevice));
        }
        // partial slots: block ids are offsets into one array per phase type
        int64_t max_spmv_blocks = 0;
        for (auto &f : fam)
            for (int i : f.members) max_spmv_blocks += ells[i]->n_chunks;
        DevBuf part((size_t)std::max<int64_t>(max_spmv_blocks, nbv) * sizeof(double) + 64);
        DevBuf part2(reorth ? (size_t)M * nbv * 2 * sizeof(double) : 64);

        auto upload_inst = [&]() {
            QMB_HIP(hipMemcpy(d_inst.p, hin.data(), n * sizeof(BInst), hipMemcpyHostToDevice));
        };
        // start vectors: per family, the vector task table; block ids must index `part`
        upload_inst();
        for (auto &f : fam) {
            if (f.vec_tasks.empty()) continue;
            // bk_start uses blockIdx.x as the partial slot: launch with the family's table, slots offset
            // by shifting the instance's vec_first to the family-local numbering
            std::vector<BInst> local = hin;
            for (size_t q = 0; q < f.members.size(); ++q) local[f.first_inst + q].vec_first -= f.vec_first_block;
            DevBuf d_local((size_t)n * sizeof(BInst));
            QMB_HIP(hipMemcpy(d_local.p, local.data(), n * sizeof(BInst), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(bk_start, dim3((unsigned)f.vec_tasks.size()), dim3(kThreads), 0, ctx->stream,
                               d_local.as<BInst>(), f.d_vec_tasks.as<BTask>(), part.as<double>(), o.seed,
                               f.cplx ? 2 : 1);
            QMB_HIP(hipGetLastError());
            hipLaunchKernelGGL(bk_finish, dim3((unsigned)f.members.size()), dim3(kThreads), 0, ctx->stream,
                               d_local.as<BInst>(), nullptr, nullptr, part.as<double>(), f.first_inst, 2);
            QMB_HIP(hipGetLastError());
            QMB_HIP(hipStreamSynchronize(ctx->stream));
        }
        // The vector-phase kernels address `part` by family-local block id; make the
        // instance descriptors family-local for the vector ranges.
        for (auto &f : fam)
            for (size_t q = 0; q < f.members.size(); ++q) hin[f.first_inst + q].vec_first -= f.vec_first_block;

        // autotune the SpMV shape per family: the four kernel variants (stream
        // policy x slots per group) at 4 chunks per block, then chunks per block
        // halved or doubled from there while that helps, under the best variant;
        // each candidate timed over 2 launches, every variant launched once
        // untimed first. Measured over all 24 shapes on five batches (6-66
        // sectors, real and complex), this search found the fastest shape each
        // time with 6-8 candidates; timing all 24, with a warm launch and 3 timed
        // ones each, took as long as the 72-sector solve itself.
        for (auto &f : fam) {
            if (f.members.empty()) continue;
            int64_t rows = 0;
            for (int i : f.members) rows += ells[i]->n_rows;
            const auto key = std::make_tuple((int)f.cplx, (int)std::log2((double)rows + 1.0), (int)f.members.size());
            TileChoice tc{};
            if (tile_cache_get(key, tc)) {
                f.cpb = tc.cpb;
                f.variant = tc.variant;
                continue;
            }
            float best = 1e30f;
            TileChoice best_c{4, 0};
            hipEvent_t ea = nullptr, eb = nullptr;
            struct TuneEvents {
                hipEvent_t *a, *b;
                ~TuneEvents() {
                    if (*a) (void)hipEventDestroy(*a);
                    if (*b) (void)hipEventDestroy(*b);
                }
            } tune_events{&ea, &eb};
            QMB_HIP(hipEventCreate(&ea));
            QMB_HIP(hipEventCreate(&eb));
            const auto &vars = spmv_variants();
            std::vector<char> warm(vars.size(), 0);
            std::map<std::pair<int, int>, float> tried;
            int cur_cpb = -1;
            DevBuf dt;
            auto measure = [&](int cpb, int vi) -> float {
                const auto it = tried.find({cpb, vi});
                if (it != tried.end()) return it->second;
                if (cpb != cur_cpb) {
                    int fb = 0;
                    build_spmv_tasks(f, cpb, fb);
                    upload_inst();
                    dt.alloc(f.spmv_tasks.size() * sizeof(BTask));
                    QMB_HIP(hipMemcpy(dt.p, f.spmv_tasks.data(), f.spmv_tasks.size() * sizeof(BTask),
                                      hipMemcpyHostToDevice));
                    cur_cpb = cpb;
                }
                const SpmvKernel k = f.cplx ? vars[vi].cplx : vars[vi].real;
                const dim3 grid((unsigned)f.spmv_tasks.size());
                if (!warm[vi]) {
                    hipLaunchKernelGGL(k, grid, dim3(kThreads), 0, ctx->stream, d_inst.as<BInst>(), dt.as<BTask>(),
                                       d_active.as<int>(), d_it.as<int>(), part.as<double>());
                    warm[vi] = 1;
                }
                QMB_HIP(hipEventRecord(ea, ctx->stream));
                for (int rep = 0; rep < 2; ++rep)
                    hipLaunchKernelGGL(k, grid, dim3(kThreads), 0, ctx->stream, d_inst.as<BInst>(), dt.as<BTask>(),
                                       d_active.as<int>(), d_it.as<int>(), part.as<double>());
                QMB_HIP(hipEventRecord(eb, ctx->stream));
                QMB_HIP(hipEventSynchronize(eb));
                float t = 0.f;
                QMB_HIP(hipEventElapsedTime(&t, ea, eb));
                tried[{cpb, vi}] = t;
                if (t < best) {
                    best = t;
                    best_c = TileChoice{cpb, vi};
                }
                return t;
            };
            for (int vi = 0; vi < (int)vars.size(); ++vi) measure(4, vi);
            const int v = best_c.variant;
            for (const int grow : {0, 1}) {
                int c = 4;
                while (true) {
                    const int nc = grow ? c * 2 : c / 2;
                    if (nc < 1 || nc > 32 || measure(nc, v) >= tried[{c, v}]) break;
                    c = nc;
                }
            }
            f.cpb = best_c.cpb;
            f.variant = best_c.variant;
            tile_cache_put(key, best_c);
        }
        // final SpMV task tables; spmv block ids are family-local too
        for (auto &f : fam) {
            if (f.members.empty()) continue;
            int fb = 0;
            build_spmv_tasks(f, f.cpb, fb);
            f.d_spmv_tasks.alloc(f.spmv_tasks.size() * sizeof(BTask));
            QMB_HIP(hipMemcpy(f.d_spmv_tasks.p, f.spmv_tasks.data(), f.spmv_tasks.size() * sizeof(BTask),
                              hipMemcpyHostToDevice));
        }
        upload_inst();
        // tuning ran real SpMVs at it = 0: restore the state the first iteration expects
        QMB_HIP(hipMemcpyAsync(d_it.p, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));

        // one iteration = every family's phases, then the counter advance
        auto enqueue_iteration = [&](hipStream_t s) {
            for (auto &f : fam) {
                if (f.members.empty()) continue;
                const unsigned nbs = (unsigned)f.spmv_tasks.size(), nvb = (unsigned)f.vec_tasks.size();
                const int vs = f.cplx ? 2 : 1;
                const unsigned nm = (unsigned)f.members.size();
                const SpmvKernel ks = f.cplx ? spmv_variants()[f.variant].cplx : spmv_variants()[f.variant].real;
                hipLaunchKernelGGL(ks, dim3(nbs), dim3(kThreads), 0, s, d_inst.as<BInst>(), f.d_spmv_tasks.as<BTask>(),
                                   d_active.as<int>(), d_it.as<int>(), part.as<double>());
                hipLaunchKernelGGL(bk_finish, dim3(nm), dim3(kThreads), 0, s, d_inst.as<BInst>(), d_active.as<int>(),
                                   d_it.as<int>(), part.as<double>(), f.first_inst, 0);
                hipLaunchKernelGGL(bk_update, dim3(nvb), dim3(kThreads), 0, s, d_inst.as<BInst>(),
                                   f.d_vec_tasks.as<BTask>(), d_active.as<int>(), d_it.as<int>(), part.as<double>(),
                                   vs);
                hipLaunchKernelGGL(bk_finish, dim3(nm), dim3(kThreads), 0, s, d_inst.as<BInst>(), d_active.as<int>(),
                                   d_it.as<int>(), part.as<double>(), f.first_inst, 1);
                if (partial)
                    hipLaunchKernelGGL(bk_pro, dim3(nm), dim3(64), 0, s, d_inst.as<BInst>(), d_active.as<int>(),
                                       d_it.as<int>(), f.first_inst, M);
                if (reorth) {
                    for (int pass = 0; pass < 2; ++pass) {
                        hipLaunchKernelGGL(f.cplx ? bk_rdots<true> : bk_rdots<false>, dim3(nvb, kJY), dim3(kThreads), 0,
                                           s, d_inst.as<BInst>(), f.d_vec_tasks.as<BTask>(), d_active.as<int>(),
                                           d_it.as<int>(), part2.as<double>(), (int)nvb, pass);
                        hipLaunchKernelGGL(bk_rreduce, dim3((unsigned)f.members.size(), 4), dim3(kThreads), 0, s,
                                           d_inst.as<BInst>(), d_active.as<int>(), d_it.as<int>(), part2.as<double>(),
                                           (int)nvb, f.first_inst, pass);
                        hipLaunchKernelGGL(f.cplx ? bk_rapply<true> : bk_rapply<false>, dim3(nvb), dim3(kThreads), 0,
                                           s, d_inst.as<BInst>(), f.d_vec_tasks.as<BTask>(), d_active.as<int>(),
                                           d_it.as<int>(), part.as<double>(), pass);
                        // pass 0 decides whether pass 1 runs, per instance, on the device
                        hipLaunchKernelGGL(bk_finish, dim3(nm), dim3(kThreads), 0, s, d_inst.as<BInst>(),
                                           d_active.as<int>(), d_it.as<int>(), part.as<double>(), f.first_inst,
                                           pass ? 4 : 3);
                    }
                }
            }
            hipLaunchKernelGGL(bk_advance, dim3(1), dim3(64), 0, s, d_it.as<int>());
        };
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        QMB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < chunk; ++q) enqueue_iteration(ctx->stream);
        QMB_HIP(hipStreamEndCapture(ctx->stream, &graph));
        QMB_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));

        // Pipelined host checks: chunk c+1 is queued before the host evaluates the
        // coefficient snapshot taken after chunk c, so the tridiagonal solves of
        // every instance overlap the GPU's next chunk. An instance that converged
        // at chunk c is dropped from chunk c+2 on (its extra steps are discarded).
        const size_t snap = (size_t)n * (3 * M + 2);
        PinnedBuf hbuf(2 * snap * sizeof(double) + 2 * (size_t)n * sizeof(int));
        double *hsnap[2] = {hbuf.as<double>(), hbuf.as<double>() + snap};
        // two staging copies of the active mask, used in turn: the copy queued after
        // chunk c may still be pending when the host next updates the mask (it waits
        // only for chunk c's snapshot event), but not two updates later
        int *hact[2] = {reinterpret_cast<int *>(hbuf.as<double>() + 2 * snap),
                        reinterpret_cast<int *>(hbuf.as<double>() + 2 * snap) + n};
        int hsel = 0;
        hipEvent_t evs[2];
        QMB_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
        QMB_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
        struct EventPair {
            hipEvent_t *e;
            ~EventPair() {
                (void)hipEventDestroy(e[0]);
                (void)hipEventDestroy(e[1]);
            }
        } ev_guard{evs};
        std::vector<std::vector<double>> ev(n, std::vector<double>(k_eigs, std::nan(""))),
            res(n, std::vector<double>(k_eigs, std::nan("")));
        std::vector<int> iters(n, 0), conv(n, 0), brk(n, 0);
        std::vector<char> done_now(n, 0);
        std::vector<std::vector<double>> zvec(vecs ? n : 0);  // lowest Ritz vector of T, per stopped instance
        int launched = 0, n_active = n;
        double ms_host = 0.0;
        // with opts.profile: HIP events around every graph replay (the GPU's busy
        // time; the rest of the wall time is set-up and host gaps)
        std::vector<hipEvent_t> rev;
        auto launch_chunk = [&](int slot) {
            if (o.profile) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                QMB_HIP(hipEventCreate(&e0));
                rev.push_back(e0);  // owned by rev_guard from here on
                QMB_HIP(hipEventCreate(&e1));
                rev.push_back(e1);
                QMB_HIP(hipEventRecord(e0, ctx->stream));
            }
            QMB_HIP(hipGraphLaunch(exec, ctx->stream));
            if (o.profile) QMB_HIP(hipEventRecord(rev.back(), ctx->stream));
            launched += chunk;
            QMB_HIP(hipMemcpyAsync(hsnap[slot], scal.p, snap * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            QMB_HIP(hipEventRecord(evs[slot], ctx->stream));
        };
        struct EventList {
            std::vector<hipEvent_t> *v;
            ~EventList() {
                for (auto e : *v) (void)hipEventDestroy(e);
            }
        } rev_guard{&rev};
        // host threads for the per-instance tridiagonal checks (independent problems)
        const int n_threads =
            (int)std::max(1u, std::min({std::thread::hardware_concurrency(), 16u, (unsigned)((n + 3) / 4)}));
        launch_chunk(0);
        int pend = 0, pend_it = launched;
        while (true) {
            const bool next = launched + chunk <= M && n_active > 0;
            if (next) launch_chunk(1 - pend);
            QMB_HIP(hipEventSynchronize(evs[pend]));
            const int it = pend_it;
            const double *hs = hsnap[pend];
            auto th0 = std::chrono::steady_clock::now();
            bool changed = false;
            // instance pos: tridiagonal solve of its snapshot, convergence test; done_now[pos] = 1 when it stops
            auto check = [&](int pos, std::vector<double> &w, std::vector<double> &zl) {
                const double *a = hs + (size_t)pos * (3 * M + 2);
                const double *b = a + M;
                double tsc = 0.0;
                for (int j = 0; j < it; ++j) tsc = std::max(tsc, std::fabs(a[j]) + b[j] + b[j + 1]);
                int m = it;
                bool bd = false;
                for (int j = 1; j <= it; ++j)
                    if (!(b[j] > 1e-13 * std::max(tsc, 1e-300))) {
                        m = j;
                        bd = true;
                        break;
                    }
                const int kk = std::min(k_eigs + 1, m);
                tridiag_lowest(m, a, b, kk, w.data(), zl.data(), nullptr);
                bool ok = bd && m >= k_eigs;
                if (!bd && m >= k_eigs) {
                    ok = true;
                    const double sq = std::sqrt(tol);
                    for (int j = 0; j < k_eigs && ok; ++j) {
                        const double sc = std::max(1.0, std::fabs(w[j]));
                        const double r = std::fabs(b[m] * zl[j]);
                        double gap = 1.0;
                        if (j + 1 < kk) gap = w[j + 1] - w[j];
                        if (j > 0) gap = std::min(gap, w[j] - w[j - 1]);
                        ok = r <= sq * sc && r * r / std::max(gap, sq) <= tol * sc;
                    }
                }
                const int kq = std::min(k_eigs, m);
                for (int j = 0; j < k_eigs; ++j) {
                    ev[pos][j] = j < kq ? w[j] : std::nan("");
                    res[pos][j] = j < kq ? std::fabs(b[m] * zl[j]) : std::nan("");
                }
                iters[pos] = m;
                done_now[pos] = 0;
                if (ok || bd || it >= M) {
                    conv[pos] = ok ? 1 : 0;
                    brk[pos] = bd ? 1 : 0;
                    done_now[pos] = 1;
                    if (vecs) {
                        zvec[pos].assign(m, 0.0);
                        tridiag_lowest(m, a, b, 1, w.data(), zl.data(), zvec[pos].data());
                    }
                }
            };
            std::vector<int> act;
            for (int pos = 0; pos < n; ++pos)
                if (active[pos]) act.push_back(pos);
            const int T = std::min<int>(n_threads, (int)act.size());
            auto run_part = [&](int t) {
                std::vector<double> w(k_eigs + 1), zl(k_eigs + 1);
                for (size_t i = t; i < act.size(); i += T) check(act[i], w, zl);
            };
            if (T <= 1) {
                run_part(0);
            } else {
                std::vector<std::thread> pool;
                for (int t = 1; t < T; ++t) pool.emplace_back(run_part, t);
                run_part(0);
                for (auto &th : pool) th.join();
            }
            for (int pos : act)
                if (done_now[pos]) {
                    active[pos] = 0;
                    --n_active;
                    changed = true;
                }
            ms_host += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
            if (!next) break;  // no chunk in flight: that snapshot was the last one
            if (changed && n_active > 0) {
                std::memcpy(hact[hsel], active.data(), n * sizeof(int));
                QMB_HIP(hipMemcpyAsync(d_active.p, hact[hsel], n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
                hsel = 1 - hsel;
            }
            pend = 1 - pend;
            pend_it = launched;
        }
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        const double wall =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
        double ms_replays = 0.0;
        for (size_t i = 0; i + 1 < rev.size(); i += 2) {
            float t = 0.f;
            QMB_HIP(hipEventElapsedTime(&t, rev[i], rev[i + 1]));
            ms_replays += t;
        }
        bool all_conv = true;
        std::vector<double> hpro((size_t)n * 3, 0.0);
        QMB_HIP(hipMemcpyAsync(hpro.data(), pros.p, hpro.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        for (int pos = 0; pos < n; ++pos) {
            const int gi = order[pos];
            for (int j = 0; j < k_eigs; ++j) evals[(size_t)gi * k_eigs + j] = ev[pos][j];
            all_conv = all_conv && conv[pos];
            if (reports) {
                qmb_lanczos_report &r = reports[gi];
                std::memset(&r, 0, sizeof(r));
                r.iters = iters[pos];
                r.converged = conv[pos];
                r.breakdown = brk[pos];
                r.n_spmv = iters[pos];
                r.n_reorth = !reorth ? 0 : partial ? (int)hpro[(size_t)pos * 3 + 2] : iters[pos];
                for (int j = 0; j < std::min(k_eigs, 8); ++j) r.residual[j] = res[pos][j];
                r.ms_total = wall;
                r.ms_host = ms_host;
                r.ms_spmv = ms_replays;
                r.n_spmv_timed = (int)(rev.size() / 2);
            }
        }
        if (vecs) {
            // ground-state vectors from the kept Krylov bases (the instances' extra
            // steps after convergence wrote only slots past their own m)
            DevBuf dz((size_t)M * sizeof(double));
            for (int pos = 0; pos < n; ++pos) {
                const int gi = order[pos];
                const qmb_ell *e = ells[gi];
                const int64_t cnt = e->n_rows * (e->value_type == QMB_COMPLEX ? 2 : 1);
                const int m = (int)zvec[pos].size();
                DevBuf X((size_t)cnt * sizeof(double));
                if (m > 0) {
                    QMB_HIP(hipMemcpyAsync(dz.p, zvec[pos].data(), (size_t)m * sizeof(double),
                                           hipMemcpyHostToDevice, ctx->stream));
                    const unsigned g = (unsigned)std::min<int64_t>(4096, (cnt + kThreads * 4 - 1) / (kThreads * 4));
                    hipLaunchKernelGGL(bk_combine, dim3(std::max(1u, g)), dim3(kThreads), 0, ctx->stream, cnt, m,
                                       hin[pos].VB, dz.as<double>(), X.as<double>());
                    QMB_HIP(hipGetLastError());
                } else {
                    QMB_HIP(hipMemsetAsync(X.p, 0, (size_t)cnt * sizeof(double), ctx->stream));
                }
                QMB_HIP(hipMemcpyAsync(vecs[gi], X.p, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost,
                                       ctx->stream));
                QMB_HIP(hipStreamSynchronize(ctx->stream));  // dz and X are reused or freed next
            }
        }
        if (!all_conv) throw Error{QMB_ERR_NOT_CONVERGED, "qmb_lanczos_batched: some instances did not converge"};
    }
}

}  // namespace qmb

extern "C" qmb_status qmb_lanczos_batched(qmb_ctx *ctx, int n, const qmb_ell *const *ells, const qmb_lanczos_opts *opts,
                                          double *evals, qmb_lanczos_report *reports) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        run_batched(ctx, n, ells, opts, evals, nullptr, reports);
    });
}

extern "C" qmb_status qmb_lanczos_batched_vectors(qmb_ctx *ctx, int n, const qmb_ell *const *ells,
                                                  const qmb_lanczos_opts *opts, double *evals, double *const *vecs,
                                                  qmb_lanczos_report *reports) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(vecs && opts, "qmb_lanczos_batched_vectors: bad argument");
        QMB_REQUIRE(opts->reorth == 1 || opts->reorth == 2,
                    "qmb_lanczos_batched_vectors: the vectors need kept Lanczos vectors (reorth = 1 or 2)");
        for (int i = 0; i < n; ++i) QMB_REQUIRE(vecs[i], "qmb_lanczos_batched_vectors: NULL vector");
        run_batched(ctx, n, ells, opts, evals, vecs, reports);
    });
}
