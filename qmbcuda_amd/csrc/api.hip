// Context, group and term-table handles, error reporting, device scans.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <numeric>
#include <unordered_map>

#include "internal.hpp"
#include "terms.hpp"

namespace qmb {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

DevPool *&current_pool() {
    static thread_local DevPool *p = nullptr;
    return p;
}

qmb_status fail(qmb_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

void ensure_binom(qmb_ctx *ctx) {
    if (ctx->binom.p) return;
    std::vector<uint64_t> t(64 * 33, 0);
    // Pascal's rule, exact in uint64 for n < 64.
    std::vector<std::vector<uint64_t>> c(65, std::vector<uint64_t>(65, 0));
    for (int n = 0; n <= 64; ++n) {
        c[n][0] = 1;
        for (int k = 1; k <= n; ++k) c[n][k] = c[n - 1][k - 1] + (k <= n - 1 ? c[n - 1][k] : 0);
    }
    for (int p = 0; p < 64; ++p)
        for (int j = 0; j < 33; ++j) t[p * 33 + j] = c[p][j];
    ctx->binom.alloc(t.size() * sizeof(uint64_t));
    QMB_HIP(hipMemcpy(ctx->binom.p, t.data(), t.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
}

// ---- exclusive scans --------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <class T>
__device__ inline T block_exclusive_scan(T v, T *lds, T &total) {
    // Hillis-Steele over 256 values in LDS; fine for tile-level use.
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        T add = (t >= off) ? lds[t - off] : T(0);
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    total = lds[kScanThreads - 1];
    T incl = lds[t];
    __syncthreads();
    return incl - v;
}

// bit words at stride 2 (the rank dictionary's records)
struct PopcountSrc {
    const uint64_t *w;
    __device__ int64_t operator()(int64_t i) const { return __popcll(w[2 * i]); }
};

struct ArraySrc {
    const int64_t *v;
    __device__ int64_t operator()(int64_t i) const { return v[i]; }
};

template <class Src>
__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(Src src, int64_t n, int64_t *tile_sums) {
    __shared__ int64_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t s = 0;
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) s += src(base + i);
    int64_t total;
    block_exclusive_scan<int64_t>(s, lds, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single-block exclusive scan of the tile sums (in place); writes the grand total
// at [n] and, when total_out is set, there too.
__global__ __launch_bounds__(kScanThreads) void scan_tile_offsets(int64_t *tile_sums, int64_t n,
                                                                  int64_t *total_out) {
    __shared__ int64_t lds[kScanThreads];
    int64_t carry = 0;
    for (int64_t base = 0; base < n; base += kScanThreads) {
        const int64_t i = base + threadIdx.x;
        const int64_t v = i < n ? tile_sums[i] : 0;
        int64_t total;
        const int64_t ex = block_exclusive_scan<int64_t>(v, lds, total);
        if (i < n) tile_sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        tile_sums[n] = carry;
        if (total_out) *total_out = carry;
    }
}

// out[i * stride + offset] = exclusive prefix of element i
template <class Src, class Out>
__global__ __launch_bounds__(kScanThreads) void scan_write(Src src, int64_t n, const int64_t *tile_off,
                                                           Out *out, int stride, int offset) {
    __shared__ int64_t lds[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t vals[kScanItems];
    int64_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        vals[i] = (base + i < n) ? src(base + i) : 0;
        s += vals[i];
    }
    int64_t total;
    int64_t run = block_exclusive_scan<int64_t>(s, lds, total) + tile_off[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[(base + i) * stride + offset] = (Out)run;
        run += vals[i];
    }
}

// Queues the three scan phases on the context's stream; returns the device
// address of the grand total (inside sums).
template <class Src, class Out>
static const int64_t *launch_scan(qmb_ctx *ctx, Src src, int64_t n, Out *out, DevBuf &sums, int64_t *total_out,
                                  int stride = 1, int offset = 0) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    sums.alloc((size_t)(tiles + 1) * sizeof(int64_t));
    if (tiles > 0) {
        hipLaunchKernelGGL(scan_tile_sums<Src>, dim3((unsigned)tiles), dim3(kScanThreads), 0, ctx->stream, src, n,
                           sums.as<int64_t>());
        QMB_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(scan_tile_offsets, dim3(1), dim3(kScanThreads), 0, ctx->stream, sums.as<int64_t>(), tiles,
                       total_out);
    QMB_HIP(hipGetLastError());
    if (tiles > 0) {
        hipLaunchKernelGGL((scan_write<Src, Out>), dim3((unsigned)tiles), dim3(kScanThreads), 0, ctx->stream, src, n,
                           sums.as<int64_t>(), out, stride, offset);
        QMB_HIP(hipGetLastError());
    }
    return sums.as<int64_t>() + tiles;
}

template <class Src, class Out>
static int64_t run_scan(qmb_ctx *ctx, Src src, int64_t n, Out *out, int stride = 1, int offset = 0) {
    DevBuf sums;
    const int64_t *d_total = launch_scan(ctx, src, n, out, sums, nullptr, stride, offset);
    int64_t *st = ctx->host_stage<int64_t>(sizeof(int64_t));
    QMB_HIP(hipMemcpyAsync(st, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    QMB_HIP(hipStreamSynchronize(ctx->stream));
    return st[0];
}

int64_t scan_popcount_dict(qmb_ctx *ctx, uint64_t *dict, int64_t n_words) {
    return run_scan(ctx, PopcountSrc{dict}, n_words, dict, 2, 1);
}

// Chunk widths (64 x the longest row of each 64-row chunk), one wavefront per
// chunk, and per-block row-length sums and maxima.
__global__ __launch_bounds__(256) void chunk_widths(const uint16_t *row_len, int64_t n_rows, int64_t n_chunks,
                                                    int64_t *width, int64_t *nnz_partial, int *max_partial) {
    __shared__ int64_t s_sum[4];
    __shared__ int s_max[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int64_t sum = 0;
    int mx = 0;
    for (int64_t c = (int64_t)blockIdx.x * 4 + wv; c < n_chunks; c += (int64_t)gridDim.x * 4) {
        const int64_t r = c * kChunk + lane;
        const int l = r < n_rows ? (int)row_len[r] : 0;
        int64_t cs = l;
        int cm = l;
        for (int off = 32; off > 0; off >>= 1) {
            cs += __shfl_xor(cs, off, 64);
            const int o = __shfl_xor(cm, off, 64);
            cm = cm > o ? cm : o;
        }
        if (lane == 0) width[c] = (int64_t)kChunk * cm;
        sum += cs;
        mx = mx > cm ? mx : cm;
    }
    if (lane == 0) {
        s_sum[wv] = sum;
        s_max[wv] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        int m = 0;
        for (int i = 0; i < 4; ++i) {
            t += s_sum[i];
            m = m > s_max[i] ? m : s_max[i];
        }
        nnz_partial[blockIdx.x] = t;
        max_partial[blockIdx.x] = m;
    }
}

// {stored, nnz, max row length, error flag} in one word group for one read-back.
__global__ __launch_bounds__(64) void chunk_stats_final(const int64_t *nnz_partial, const int *max_partial,
                                                        int blocks, const int64_t *total, const int *err,
                                                        int64_t *out) {
    int64_t t = 0;
    int m = 0;
    for (int i = threadIdx.x; i < blocks; i += 64) {
        t += nnz_partial[i];
        m = m > max_partial[i] ? m : max_partial[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        t += __shfl_xor(t, off, 64);
        const int o = __shfl_xor(m, off, 64);
        m = m > o ? m : o;
    }
    if (threadIdx.x != 0) return;
    out[0] = *total;
    out[1] = t;
    out[2] = m;
    out[3] = err ? *err : 0;
}

int64_t scan_chunks_i64(qmb_ctx *ctx, const uint16_t *row_len, int64_t n_rows, int64_t *chunk_ptr,
                        int64_t n_chunks, int64_t *nnz_out, int *max_len_out, const int *d_err, int *err_out) {
    const int blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n_chunks + 3) / 4));
    DevBuf wid((size_t)std::max<int64_t>(n_chunks, 1) * sizeof(int64_t)), np((size_t)blocks * sizeof(int64_t)),
        mp((size_t)blocks * sizeof(int)), stats(4 * sizeof(int64_t)), sums;
    hipLaunchKernelGGL(chunk_widths, dim3(blocks), dim3(256), 0, ctx->stream, row_len, n_rows, n_chunks,
                       wid.as<int64_t>(), np.as<int64_t>(), mp.as<int>());
    QMB_HIP(hipGetLastError());
    const int64_t *d_total = launch_scan(ctx, ArraySrc{wid.as<int64_t>()}, n_chunks, chunk_ptr, sums,
                                         chunk_ptr + n_chunks);
    hipLaunchKernelGGL(chunk_stats_final, dim3(1), dim3(64), 0, ctx->stream, np.as<int64_t>(), mp.as<int>(),
                       blocks, d_total, d_err, stats.as<int64_t>());
    QMB_HIP(hipGetLastError());
    int64_t *h = ctx->host_stage<int64_t>(4 * sizeof(int64_t));
    QMB_HIP(hipMemcpyAsync(h, stats.p, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    QMB_HIP(hipStreamSynchronize(ctx->stream));
    *nnz_out = h[1];
    *max_len_out = (int)h[2];
    if (err_out) *err_out = (int)h[3];
    return h[0];
}

// ---- group analysis ---------------------------------------------------------
// Finds a block width B for which every element is a cyclic translation:
// site p -> ((p/B + a) mod (N/B))*B + ((p mod B) + s) mod B.
static bool detect_torus(qmb_group &g) {
    const int N = g.n_sites;
    for (int B = N; B >= 1; --B) {
        if (N % B) continue;
        const int nb = N / B;
        std::vector<uint64_t> keep(g.G);
        std::vector<uint8_t> sm(g.G), bg(g.G);
        bool ok = true;
        for (int e = 0; e < g.G && ok; ++e) {
            const uint32_t *p = &g.perm[(size_t)e * N];
            const int a = (int)(p[0] / B), s = (int)(p[0] % B);
            for (int q = 0; q < N && ok; ++q) {
                const int want = (((q / B) + a) % nb) * B + ((q % B) + s) % B;
                if ((int)p[q] != want) ok = false;
            }
            if (!ok) break;
            uint64_t k = 0;
            for (int blk = 0; blk < nb; ++blk)
                for (int j = 0; j < B - s; ++j) k |= 1ULL << (blk * B + j);
            keep[e] = k;
            sm[e] = (uint8_t)s;
            bg[e] = (uint8_t)(a * B);
        }
        if (ok) {
            g.torus = 1;
            g.block = B;
            g.keep = keep;
            g.small = sm;
            g.big = bg;
            return true;
        }
    }
    g.torus = 0;
    g.block = N;
    g.keep.assign(g.G, 0);
    g.small.assign(g.G, 0);
    g.big.assign(g.G, 0);
    return false;
}

// Orbit-minimum tables for a group that is the whole product of block shifts
// a < nb, in-block shifts s < B (and, with flip, both halves), B <= 8 and
// nb <= 8: the top block of an image is one source block rotated (and maybe
// complemented), so the minimum's top block and the elements that can reach it
// follow from per-block-value tables instead of all G images.
//   tab[v]        = min over s of rot_B(v, s)          (v < 2^B)
//   tab[256 + v]  = bit s set when rot_B(v, s) equals it
//   tab[512 + (f*8 + a)*8 + s] = element index of (flip f, block shift a, s)
// Rings (one block of N sites, every rotation s and, with flip, both halves,
// 8 <= N <= 56): the top 8 bits of an image are a circular 8-site window of
// the source, so only the rotations whose window is the smallest need full
// images. tab[512 + f*64 + s] = element index of (flip f, rotation s).
static bool build_ring_lut(qmb_group &g) {
    const int N = g.n_sites;
    if (g.block != N || N < 8 || N > 56) return false;
    int hf = 0;
    for (int e = 0; e < g.G; ++e) hf |= g.flip[e] ? 1 : 0;
    std::vector<int> gid(2 * 64, -1);
    for (int e = 0; e < g.G; ++e) {
        if (g.big[e] != 0) return false;
        const int i = (g.flip[e] ? 1 : 0) * 64 + g.small[e];
        if (gid[i] >= 0) return false;
        gid[i] = e;
    }
    for (int f = 0; f <= hf; ++f)
        for (int s = 0; s < N; ++s)
            if (gid[f * 64 + s] < 0) return false;
    g.lut_tab.assign(kLutBytes, 0);
    for (int i = 0; i < 2 * 64; ++i) g.lut_tab[512 + i] = (uint8_t)(gid[i] < 0 ? 0 : gid[i]);
    g.lut = 2;
    g.lut_flip = hf;
    return true;
}

// A group that is not a ring itself but holds every rotation of the ring (with
// or without the spin flip) as a subgroup H, e.g. translations x spin flip x a
// reflection: G = H c_0 + H c_1 + ... (c_0 = identity, at most kMaxCosets
// cosets). The orbit minimum is then the smallest of the ring minima of the
// images c_i x (lut = 3); cos_tab names the element (rotation s, flip f) o c_i.
static void build_coset_ring(qmb_group &g) {
    const int N = g.n_sites, G = g.G;
    if (g.torus || N < 8 || N > 56) return;
    std::vector<int> gid(128, -1), in_h(G, 0);
    for (int e = 0; e < G; ++e) {
        const uint32_t *p = &g.perm[(size_t)e * N];
        const int sh = (int)p[0];
        bool rot = true;
        for (int q = 0; q < N && rot; ++q) rot = (int)p[q] == (q + sh) % N;
        if (!rot) continue;
        const int i = (g.flip[e] ? 64 : 0) + sh;
        if (gid[i] >= 0) return;  // repeated element
        gid[i] = e;
        in_h[e] = 1;
    }
    int hf = 0;
    for (int e = 0; e < G; ++e) hf |= (in_h[e] && g.flip[e]) ? 1 : 0;
    int nh = 0;
    for (int f = 0; f <= hf; ++f)
        for (int sh = 0; sh < N; ++sh) {
            if (gid[f * 64 + sh] < 0) return;
            ++nh;
        }
    if (nh == G || G % nh || G / nh > kMaxCosets || gid[0] != 0) return;
    // cosets H c: c = the first element not yet covered
    std::vector<int> covered(G, 0), reps;
    for (int e = 0; e < G; ++e) {
        if (covered[e]) continue;
        reps.push_back(e);
        for (int h = 0; h < G; ++h)
            if (in_h[h]) covered[g.comp[(size_t)h * G + e]] = 1;
    }
    if ((int)reps.size() * nh != G) return;
    // a coset holding the site reversal s -> N-1-s (no flip) is represented by
    // it: its image is one bit reversal on the device
    g.cos_rev = 0;
    for (size_t c = 1; c < reps.size(); ++c)
        for (int h = 0; h < G; ++h) {
            if (!in_h[h]) continue;
            const int e = g.comp[(size_t)h * G + reps[c]];
            bool rev = !g.flip[e];
            for (int q = 0; q < N && rev; ++q) rev = (int)g.perm[(size_t)e * N + q] == N - 1 - q;
            if (rev) {
                reps[c] = e;
                g.cos_rev |= 1 << c;
                break;
            }
        }
    g.cos_tab.assign(kCosBytes, 0);
    for (size_t c = 0; c < reps.size(); ++c)
        for (int i = 0; i < 128; ++i)
            if (gid[i] >= 0) g.cos_tab[c * 128 + i] = g.comp[(size_t)gid[i] * G + reps[c]];
    g.n_cos = (int)reps.size();
    g.lut = 3;
    g.lut_flip = hf;
}

static void build_lut(qmb_group &g) {
    g.lut = 0;
    g.lut_flip = 0;
    g.lut_tab.clear();
    if (!g.torus) return;
    if (build_ring_lut(g)) return;
    const int N = g.n_sites, B = g.block, nb = N / B;
    if (B > 8 || nb > 8) return;
    int hf = 0;
    for (int e = 0; e < g.G; ++e) hf |= g.flip[e] ? 1 : 0;
    std::vector<int> gid(2 * 8 * 8, -1);
    for (int e = 0; e < g.G; ++e) {
        const int a = g.big[e] / B, s = g.small[e], f = g.flip[e] ? 1 : 0;
        const int i = (f * 8 + a) * 8 + s;
        if (gid[i] >= 0) return;
        gid[i] = e;
    }
    for (int f = 0; f <= hf; ++f)
        for (int a = 0; a < nb; ++a)
            for (int s = 0; s < B; ++s)
                if (gid[(f * 8 + a) * 8 + s] < 0) return;
    g.lut_tab.assign(kLutBytes, 0);
    const unsigned mB = (1u << B) - 1u;
    for (unsigned v = 0; v <= mB; ++v) {
        unsigned best = 0xFF, mask = 0;
        for (int s = 0; s < B; ++s) {
            const unsigned r = s ? (((v << s) | (v >> (B - s))) & mB) : v;
            if (r < best) {
                best = r;
                mask = 1u << s;
            } else if (r == best) {
                mask |= 1u << s;
            }
        }
        g.lut_tab[v] = (uint8_t)best;
        g.lut_tab[256 + v] = (uint8_t)mask;
    }
    for (int i = 0; i < 2 * 8 * 8; ++i) g.lut_tab[512 + i] = (uint8_t)(gid[i] < 0 ? 0 : gid[i]);
    g.lut = 1;
    g.lut_flip = hf;
}

}  // namespace qmb

using namespace qmb;

extern "C" {

const char *qmb_last_error(void) { return qmb::g_last_error.c_str(); }

int qmb_version(void) { return 10000; }

int qmb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

qmb_status qmb_ctx_create(int device, qmb_ctx **out) {
    return guarded([&] {
        QMB_REQUIRE(out != nullptr, "qmb_ctx_create: out is NULL");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            throw Error{QMB_ERR_NO_DEVICE, "qmb_ctx_create: no HIP device"};
        QMB_REQUIRE(device >= 0 && device < n, "qmb_ctx_create: device index out of range");
        QMB_HIP(hipSetDevice(device));
        auto *c = new qmb_ctx();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            c->stream = nullptr;
            delete c;
            throw Error{QMB_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        c->pool.stream = c->stream;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
            c->pool.max_cached = std::max<size_t>(c->pool.max_cached, total_b / 4 * 3);
        else
            (void)hipGetLastError();
        try {
            ensure_binom(c);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

qmb_status qmb_ctx_destroy(qmb_ctx *ctx) {
    if (!ctx) return QMB_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    qmb::release(ctx);  // torn down once no basis, ELL or network uses it
    return QMB_OK;
}

qmb_status qmb_ctx_synchronize(qmb_ctx *ctx) {
    return guarded([&] {
        QMB_REQUIRE(ctx, "qmb_ctx_synchronize: NULL context");
        QMB_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// Closure under composition, and which character rows are one-dimensional
// representations (chi(a b) = chi(a) chi(b) for every pair, equal characters on
// repeated elements). Point operations combined with translations give rows
// that are not (momenta the operation does not fix); those rows stay in the
// table but no basis can be built for them.
static void check_closure_and_irreps(qmb_group &g) {
    const int n = g.n_sites, G = g.G;
    std::unordered_map<std::string, int> index;
    auto key = [&](const uint32_t *p, int f) {
        std::string k((size_t)n + 1, '\0');
        for (int s = 0; s < n; ++s) k[s] = (char)p[s];
        k[n] = (char)f;
        return k;
    };
    std::vector<int> first(G);
    for (int e = 0; e < G; ++e) first[e] = index.emplace(key(&g.perm[(size_t)e * n], g.flip[e]), e).first->second;
    g.irrep_ok.assign(g.n_irreps, 1);
    auto chi = [&](int a, int e) { return std::complex<double>(g.chars[((size_t)a * G + e) * 2], g.chars[((size_t)a * G + e) * 2 + 1]); };
    const double tol = 1e-9;
    for (int e = 0; e < G; ++e)
        for (int a = 0; a < g.n_irreps; ++a)
            if (std::abs(chi(a, e) - chi(a, first[e])) > tol) g.irrep_ok[a] = 0;
    std::vector<uint32_t> pr(n);
    g.comp.assign((size_t)G * G, 0);
    for (int x = 0; x < G; ++x)
        for (int y = 0; y < G; ++y) {
            const uint32_t *px = &g.perm[(size_t)x * n], *py = &g.perm[(size_t)y * n];
            for (int s = 0; s < n; ++s) pr[s] = px[py[s]];
            auto it = index.find(key(pr.data(), g.flip[x] ^ g.flip[y]));
            QMB_REQUIRE(it != index.end(), "qmb_group_create: element list is not closed under composition");
            const int z = it->second;
            g.comp[(size_t)x * G + y] = (uint8_t)z;
            for (int a = 0; a < g.n_irreps; ++a)
                if (g.irrep_ok[a] && std::abs(chi(a, z) - chi(a, x) * chi(a, y)) > tol) g.irrep_ok[a] = 0;
        }
}

qmb_status qmb_group_create(int n_sites, int n_elems, const uint32_t *group_el, const uint8_t *flip,
                            int n_irreps, const double *chars, qmb_group **out) {
    return guarded([&] {
        QMB_REQUIRE(out, "qmb_group_create: out is NULL");
        *out = nullptr;
        QMB_REQUIRE(n_sites >= 1 && n_sites <= 64, "qmb_group_create: n_sites must be in [1, 64]");
        QMB_REQUIRE(n_elems >= 1 && n_elems <= kMaxGroup, "qmb_group_create: group size must be in [1, 128]");
        QMB_REQUIRE(n_irreps >= 1, "qmb_group_create: need at least one irrep");
        QMB_REQUIRE(group_el && chars, "qmb_group_create: NULL table");
        auto g = std::make_unique<qmb_group>();
        g->n_sites = n_sites;
        g->G = n_elems;
        g->n_irreps = n_irreps;
        g->perm.assign(group_el, group_el + (size_t)n_elems * n_sites);
        g->flip.assign(n_elems, 0);
        if (flip)
            for (int e = 0; e < n_elems; ++e) g->flip[e] = flip[e] ? 1 : 0;
        g->chars.assign(chars, chars + (size_t)n_irreps * n_elems * 2);
        for (int e = 0; e < n_elems; ++e) {
            std::vector<int> seen(n_sites, 0);
            for (int s = 0; s < n_sites; ++s) {
                const uint32_t d = g->perm[(size_t)e * n_sites + s];
                QMB_REQUIRE(d < (uint32_t)n_sites, "qmb_group_create: destination site out of range");
                QMB_REQUIRE(!seen[d], "qmb_group_create: element is not a permutation");
                seen[d] = 1;
            }
        }
        for (int s = 0; s < n_sites; ++s)
            QMB_REQUIRE(g->perm[s] == (uint32_t)s && !g->flip[0], "qmb_group_create: element 0 must be the identity");
        // inverse table: h with perm_h[perm_g[s]] == s and equal flip
        g->inv.assign(n_elems, -1);
        for (int e = 0; e < n_elems; ++e) {
            for (int h = 0; h < n_elems && g->inv[e] < 0; ++h) {
                if (g->flip[h] != g->flip[e]) continue;
                bool ok = true;
                for (int s = 0; s < n_sites && ok; ++s)
                    ok = g->perm[(size_t)h * n_sites + g->perm[(size_t)e * n_sites + s]] == (uint32_t)s;
                if (ok) g->inv[e] = h;
            }
            QMB_REQUIRE(g->inv[e] >= 0, "qmb_group_create: element list is not closed under inversion");
        }
        check_closure_and_irreps(*g);
        uint64_t hsh = 1469598103934665603ULL;  // FNV-1a, identifies the group in saved files
        auto mix = [&](const void *p, size_t n) {
            const unsigned char *c = static_cast<const unsigned char *>(p);
            for (size_t i = 0; i < n; ++i) hsh = (hsh ^ c[i]) * 1099511628211ULL;
        };
        mix(g->perm.data(), g->perm.size() * sizeof(uint32_t));
        mix(g->flip.data(), g->flip.size());
        mix(g->chars.data(), g->chars.size() * sizeof(double));
        g->hash = hsh;
        g->orbit_norms = orbit_norm_table(n_elems);
        detect_torus(*g);
        build_lut(*g);
        if (!g->lut) build_coset_ring(*g);
        *out = g.release();
    });
}

qmb_status qmb_group_irreps_valid(const qmb_group *g, int32_t *valid) {
    return guarded([&] {
        QMB_REQUIRE(g && valid, "qmb_group_irreps_valid: NULL argument");
        for (int a = 0; a < g->n_irreps; ++a) valid[a] = g->irrep_ok[a];
    });
}

qmb_status qmb_group_destroy(qmb_group *g) {
    qmb::release(g);
    return QMB_OK;
}

qmb_status qmb_group_info(const qmb_group *g, int32_t *inv, int *torus_path) {
    return guarded([&] {
        QMB_REQUIRE(g, "qmb_group_info: NULL group");
        if (inv) std::copy(g->inv.begin(), g->inv.end(), inv);
        if (torus_path) *torus_path = g->torus;
    });
}

qmb_status qmb_terms_create(int n_sites, int n_terms, int max_ops, const int32_t *op_type, const int32_t *site,
                            const double *scalar, const int32_t *n_ops, qmb_terms **out) {
    return guarded([&] {
        QMB_REQUIRE(out, "qmb_terms_create: out is NULL");
        *out = nullptr;
        QMB_REQUIRE(n_sites >= 1 && n_sites <= 64, "qmb_terms_create: n_sites must be in [1, 64]");
        QMB_REQUIRE(n_terms >= 0 && max_ops >= 0 && max_ops <= 16, "qmb_terms_create: bad table shape");
        QMB_REQUIRE(n_terms == 0 || (op_type && site && scalar && n_ops), "qmb_terms_create: NULL table");
        auto t = std::make_unique<qmb_terms>();
        t->n_sites = n_sites;
        t->n_terms = n_terms;
        t->max_ops = max_ops;
        t->op_type.assign(op_type, op_type + (size_t)n_terms * max_ops);
        t->site.assign(site, site + (size_t)n_terms * max_ops);
        t->scalar.assign(scalar, scalar + (size_t)n_terms * 2);
        t->n_ops.assign(n_ops, n_ops + n_terms);
        for (int h = 0; h < n_terms; ++h) {
            QMB_REQUIRE(n_ops[h] >= 0 && n_ops[h] <= max_ops, "qmb_terms_create: n_ops out of range");
            int up = 0, dn = 0;
            for (int k = 0; k < n_ops[h]; ++k) {
                const int ty = op_type[(size_t)h * max_ops + k];
                const int s = site[(size_t)h * max_ops + k];
                if (ty != QMB_OP_SZ && ty != QMB_OP_SP && ty != QMB_OP_SM)
                    throw Error{QMB_ERR_UNSUPPORTED, "qmb_terms_create: only Sz, Sp and Sm operators are supported"};
                QMB_REQUIRE(s >= 0 && s < n_sites, "qmb_terms_create: operator site out of range");
                up += ty == QMB_OP_SP;
                dn += ty == QMB_OP_SM;
            }
            QMB_REQUIRE(up == dn, "qmb_terms_create: term " + std::to_string(h) + " does not conserve Sz");
        }
        t->compiled.resize((size_t)n_terms * sizeof(TermC));
        for (int h = 0; h < n_terms; ++h) {
            const TermC c = compile_term(*t, h);
            std::memcpy(t->compiled.data() + (size_t)h * sizeof(TermC), &c, sizeof(TermC));
        }
        *out = t.release();
    });
}

qmb_status qmb_terms_destroy(qmb_terms *t) {
    delete t;
    return QMB_OK;
}

}  // extern "C"
