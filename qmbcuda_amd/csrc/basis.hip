// Symmetry-reduced basis on device.
//
// Replaces the reference's host Sz enumeration plus its five full-sector
// kernels (BuildBasisStates, SetUpSRS, SolveSRS_min, sort/unique,
// SetUpSRS_from_indices, ComputeSRSNorms, ComputeOrbitIndices:
// Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:277-400,
// Q/exact_diagonalization/utils.cu:39-80,716-805). Nothing of size
// C(N,k)*G is materialised:
//   mark    : one thread per 64 consecutive sector ranks; unrank the first,
//             step with the next-combination trick, keep a state when it is
//             the minimum of its orbit (and, in the compact layout, when the
//             irrep's character is 1 on its stabiliser) -> one bit word;
//   scan    : exclusive popcount scan of the words -> row offsets, stored
//             beside each word (16-byte records);
//   compact : rewrite the kept states in ascending order with their
//             stabiliser size and null flag.
// The records also serve as the rank -> row dictionary of the ELL build
// (replacing the C(N,k)*2 orbit_indices table): one random 16-byte read per
// lookup.
#include <cmath>
#include <cstring>
#include <memory>

#include "device_common.hpp"

qmb::GroupView qmb_basis::view() const {
    qmb::GroupView v;
    v.n_sites = group->n_sites;
    v.G = group->G;
    v.torus = group->torus;
    v.block = group->block;
    v.full = group->n_sites == 64 ? ~0ULL : ((1ULL << group->n_sites) - 1ULL);
    v.rep_unit = 0;
    if (group->torus)
        for (int b0 = 0; b0 < group->n_sites; b0 += group->block) v.rep_unit |= 1ULL << b0;
    const uint8_t *t = g_tab.as<uint8_t>();
    v.keep = reinterpret_cast<const uint64_t *>(t);
    v.small = t + o_small;
    v.big = t + o_big;
    v.flip = t + o_flip;
    v.perm = t + o_perm;
    v.inv = t + o_inv;
    v.lut = group->lut;
    v.lut_flip = group->lut_flip;
    v.lut_tab = (group->lut == 1 || group->lut == 2) ? t + o_lut : nullptr;
    v.n_cos = group->lut == 3 ? group->n_cos : 0;
    v.cos_rev = group->lut == 3 ? group->cos_rev : 0;
    v.cos_tab = group->lut == 3 ? t + o_cos : nullptr;
    return v;
}

namespace qmb {

// 8 waves per block share the staged binomial and group tables (about 28 KB of
// LDS), so more waves fit a CU than with 4-wave blocks (6x6 sector basis: 43.5
// -> 35 ms)
constexpr int kBasisThreads = 512;

__device__ inline bool compatible(uint64_t slo, uint64_t shi, uint64_t tlo, uint64_t thi) {
    return ((slo & ~tlo) | (shi & ~thi)) == 0;
}

template <class S>
__global__ __launch_bounds__(kBasisThreads) void basis_mark(GroupView gv, const uint64_t *binom, int n_up,
                                                            int64_t dim, int64_t n_words, int compact,
                                                            uint64_t triv_lo, uint64_t triv_hi,
                                                            uint64_t *dict) {
    __shared__ uint8_t perm_lds[kMaxGroup * 64];
    __shared__ uint64_t bin[64 * kBinJ];
    __shared__ GroupLds gt;
    stage_perm(perm_lds, gv);
    stage_binom(bin, binom);
    stage_group(gt, gv);
    __syncthreads();
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    const int64_t r0 = w * 64;
    const int cnt = (int)((dim - r0) < 64 ? (dim - r0) : 64);
    S x = colex_unrank<S>(r0, n_up, gv.n_sites, bin);
    uint64_t bits = 0;
    if (gv.lut == 2) {
        // Ring group, two passes. Pass 1 keeps the states that are not above
        // their rotations by 1, 2 and 3 sites (nor, with the spin flip, above
        // their complements or theirs): a necessary condition, branch-free.
        // Pass 2 takes each lane's survivors in turn: the rotations by
        // s = 1 .. N-1 one bit at a time (and their complements), exit at the
        // first smaller image; element indices are looked up only for images
        // equal to the state (stabiliser).
        const int N = gv.n_sites;
        const S full = (S)gv.full;
        const bool hf = gv.lut_flip != 0;
        auto rotl = [&](S v, int k) { return (S)(((v << k) | (v >> (N - k))) & full); };
        uint64_t surv = 0;
        S y = x;
        for (int i = 0; i < cnt; ++i) {
            const S a1 = rotl(y, 1), a2 = rotl(y, 2), a3 = rotl(y, 3);
            bool ok = a1 >= y && a2 >= y && a3 >= y;
            if (hf) ok = ok && (S)(y ^ full) >= y && (S)(a1 ^ full) >= y && (S)(a2 ^ full) >= y && (S)(a3 ^ full) >= y;
            if (ok) surv |= 1ULL << i;
            if (i + 1 < cnt) y = next_comb<S>(y);
        }
        y = x;
        int at = 0;
        while (surv) {
            const int i = __builtin_ctzll(surv);
            surv &= surv - 1;
            for (; at < i; ++at) y = next_comb<S>(y);
            uint64_t slo = 1, shi = 0;  // element 0 (identity) fixes y
            auto mark = [&](int f, int sh) {
                const int g = gt.lut[512 + f * 64 + sh];
                if (g < 64)
                    slo |= 1ULL << g;
                else
                    shi |= 1ULL << (g - 64);
            };
            bool rep = true;
            S z = y;
            for (int st = 0; st < N && rep; ++st) {
                if (st > 0) {
                    z = rotl(z, 1);
                    if (z < y)
                        rep = false;
                    else if (z == y)
                        mark(0, st);
                }
                if (rep && hf) {
                    const S zf = (S)(z ^ full);
                    if (zf < y)
                        rep = false;
                    else if (zf == y)
                        mark(1, st);
                }
            }
            if (rep && (!compact || compatible(slo, shi, triv_lo, triv_hi))) bits |= 1ULL << i;
        }
        dict[2 * w] = bits;
        return;
    }
    if (gv.lut == 1) {
        // Full torus product, two passes. Pass 1 keeps the states whose top block
        // equals the smallest in-block rotation of any block (complemented too
        // under flip): a necessary condition. The top block t is at most its own
        // rotation minimum only when it is one (lut[t] <= t), so the condition
        // is t == lut[t], t <= lut[t ^ mB] under flip, and t <= lut[v], lut[v ^ mB]
        // for the other blocks; the first part rejects most states with two
        // lookups, and since the top block changes rarely in colex order, whole
        // waves take that exit together. Pass 2 takes each lane's survivors in
        // turn (stepping its state forward) and decides them from the candidate
        // images alone.
        const int B = gv.block, nb = gv.n_sites / B;
        const unsigned mB = (1u << B) - 1u;
        const int topsh = (nb - 1) * B;
        const bool hf = gv.lut_flip != 0;
        uint64_t surv = 0;
        S y = x;
        for (int i = 0; i < cnt; ++i) {
            const unsigned t = (unsigned)(y >> topsh) & mB;
            bool ok = gt.lut[t] == t && (!hf || gt.lut[t ^ mB] >= t);
            if (ok) {
                for (int j = 0; j < nb - 1; ++j) {
                    const unsigned v = (unsigned)(y >> (j * B)) & mB;
                    ok = ok && gt.lut[v] >= t && (!hf || gt.lut[v ^ mB] >= t);
                }
            }
            if (ok) surv |= 1ULL << i;
            if (i + 1 < cnt) y = next_comb<S>(y);
        }
        y = x;
        int at = 0;
        while (surv) {
            const int i = __builtin_ctzll(surv);
            surv &= surv - 1;
            for (; at < i; ++at) y = next_comb<S>(y);
            uint64_t slo, shi;
            if (rep_stab_lut<S>(y, gv, gt, slo, shi) && (!compact || compatible(slo, shi, triv_lo, triv_hi)))
                bits |= 1ULL << i;
        }
        dict[2 * w] = bits;
        return;
    }
    if (gv.lut == 3) {
        // Ring rotations times cosets, two passes as for the ring: the same
        // branch-free filter (rotations by 1-3 sites and complements, images
        // under the ring subgroup), then each lane's survivors: per coset c_i
        // one image y = c_i x, then its rotations one bit at a time, exit at
        // the first smaller image.
        const int N = gv.n_sites;
        const S full = (S)gv.full;
        const bool hf = gv.lut_flip != 0;
        auto rotl = [&](S v, int k) { return (S)(((v << k) | (v >> (N - k))) & full); };
        uint64_t surv = 0;
        S y = x;
        for (int i = 0; i < cnt; ++i) {
            const S a1 = rotl(y, 1), a2 = rotl(y, 2), a3 = rotl(y, 3);
            bool ok = a1 >= y && a2 >= y && a3 >= y;
            if (hf) ok = ok && (S)(y ^ full) >= y && (S)(a1 ^ full) >= y && (S)(a2 ^ full) >= y && (S)(a3 ^ full) >= y;
            if (ok) surv |= 1ULL << i;
            if (i + 1 < cnt) y = next_comb<S>(y);
        }
        y = x;
        int at = 0;
        while (surv) {
            const int i = __builtin_ctzll(surv);
            surv &= surv - 1;
            for (; at < i; ++at) y = next_comb<S>(y);
            uint64_t slo = 0, shi = 0;
            auto mark = [&](int g) {
                if (g < 64)
                    slo |= 1ULL << g;
                else
                    shi |= 1ULL << (g - 64);
            };
            bool rep = true;
            for (int ci = 0; ci < gv.n_cos && rep; ++ci) {
                const uint8_t *ct = gt.cos + ci * 128;
                S z = ci ? coset_image<S>(y, ci, gv, gt, perm_lds) : y;
                for (int st = 0; st < N && rep; ++st) {
                    if (st > 0) z = rotl(z, 1);
                    if (z < y)
                        rep = false;
                    else if (z == y)
                        mark(ct[st]);
                    if (rep && hf) {
                        const S zf = (S)(z ^ full);
                        if (zf < y)
                            rep = false;
                        else if (zf == y)
                            mark(ct[64 + st]);
                    }
                }
            }
            if (rep && (!compact || compatible(slo, shi, triv_lo, triv_hi))) bits |= 1ULL << i;
        }
        dict[2 * w] = bits;
        return;
    }
    // One flat loop over (state, element) steps per lane: a lane moves to its
    // next state as soon as an image is smaller (most states exit after a few
    // images), so lanes never wait for a neighbour's full orbit scan.
    int i = 0, g = 1;
    uint64_t slo = 1, shi = 0;  // element 0 (identity) always fixes x
    while (i < cnt) {
        bool done = false, rep = false;
        if (g >= gv.G) {
            done = true;
            rep = true;
        } else {
            const S y = apply_elem_div<S>(x, g, gv, gt, perm_lds);
            if (y < x) {
                done = true;
            } else {
                if (y == x) {
                    if (g < 64)
                        slo |= 1ULL << g;
                    else
                        shi |= 1ULL << (g - 64);
                }
                ++g;
            }
        }
        if (done) {
            if (rep && (!compact || compatible(slo, shi, triv_lo, triv_hi))) bits |= 1ULL << i;
            ++i;
            if (i < cnt) x = next_comb<S>(x);
            g = 1;
            slo = 1;
            shi = 0;
        }
    }
    dict[2 * w] = bits;
}

// One thread per row: the row's bit word is the last word whose row offset
// is <= row (binary search over the offsets), its bit the (row - offset)-th
// set bit of that word. Every thread does the same work (one unrank, one
// stabiliser count) and the writes are coalesced.
template <class S>
__global__ __launch_bounds__(kBasisThreads) void basis_compact(GroupView gv, const uint64_t *binom, int n_up,
                                                               int64_t n_rows, int64_t n_words,
                                                               uint64_t triv_lo, uint64_t triv_hi,
                                                               const uint64_t *dict,
                                                               uint64_t *reps, uint8_t *stab, uint8_t *nullf) {
    __shared__ uint8_t perm_lds[kMaxGroup * 64];
    __shared__ uint64_t bin[64 * kBinJ];
    __shared__ GroupLds gt;
    stage_perm(perm_lds, gv);
    stage_binom(bin, binom);
    stage_group(gt, gv);
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int64_t lo = 0, hi = n_words - 1;  // the first word has 0 rows before it
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if ((int64_t)dict[2 * mid + 1] <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    uint64_t bits = dict[2 * lo];
    for (int64_t k = r - (int64_t)dict[2 * lo + 1]; k > 0; --k) bits &= bits - 1;
    const int i = __builtin_ctzll(bits);
    const S x = colex_unrank<S>(lo * 64 + i, n_up, gv.n_sites, bin);
    uint64_t slo, shi;
    if (gv.lut == 1)
        rep_stab_lut<S>(x, gv, gt, slo, shi);  // x is a representative: the candidates hold its stabiliser
    else
        is_representative<S>(x, gv, perm_lds, slo, shi);
    reps[r] = (uint64_t)x;
    stab[r] = (uint8_t)(__popcll(slo) + __popcll(shi));
    nullf[r] = compatible(slo, shi, triv_lo, triv_hi) ? 0 : 1;
}

template <class S>
__global__ __launch_bounds__(kBasisThreads) void basis_lookup(GroupView gv, const uint64_t *binom, int64_t n,
                                                              const uint64_t *states, const uint64_t *dict,
                                                              int64_t *rows,
                                                              int32_t *elems) {
    __shared__ uint8_t perm_lds[kMaxGroup * 64];
    __shared__ uint64_t bin[64 * kBinJ];
    stage_perm(perm_lds, gv);
    stage_binom(bin, binom);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int gm, mult;
    const S rep = orbit_min<S>((S)states[i], gv, perm_lds, gm, mult);
    rows[i] = row_of_rep<S>(rep, bin, dict);
    elems[i] = gm;
}

static void upload_group(qmb_basis *b, const qmb_group *g) {
    const size_t G = (size_t)g->G, n = (size_t)g->n_sites;
    auto up16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    b->o_small = up16(G * 8);
    b->o_big = b->o_small + G;
    b->o_flip = b->o_big + G;
    b->o_inv = b->o_flip + G;
    b->o_perm = up16(b->o_inv + G);
    b->o_lut = up16(b->o_perm + G * n);
    b->o_cos = b->o_lut + kLutBytes;
    const size_t total = up16(b->o_cos + kCosBytes);
    std::vector<uint8_t> h(total, 0);
    std::memcpy(h.data(), g->keep.data(), G * 8);
    for (size_t e = 0; e < G; ++e) {
        h[b->o_small + e] = g->small[e];
        h[b->o_big + e] = g->big[e];
        h[b->o_flip + e] = g->flip[e];
        h[b->o_inv + e] = (uint8_t)g->inv[e];
    }
    for (size_t i = 0; i < G * n; ++i) h[b->o_perm + i] = (uint8_t)g->perm[i];
    if (g->lut == 1 || g->lut == 2) std::memcpy(h.data() + b->o_lut, g->lut_tab.data(), kLutBytes);
    if (g->lut == 3) std::memcpy(h.data() + b->o_cos, g->cos_tab.data(), kCosBytes);
    b->g_tab.alloc(total);
    QMB_HIP(hipMemcpy(b->g_tab.p, h.data(), total, hipMemcpyHostToDevice));
}

static uint64_t binom_host(int n, int k) {
    if (k < 0 || k > n) return 0;
    uint64_t acc = 1;
    for (int i = 1; i <= k; ++i) {
        // exact: acc * (n-k+i) / i stays integral at every step
        unsigned __int128 t = (unsigned __int128)acc * (uint64_t)(n - k + i);
        acc = (uint64_t)(t / (uint64_t)i);
    }
    return acc;
}

}  // namespace qmb

using namespace qmb;

extern "C" {

qmb_status qmb_basis_build(qmb_ctx *ctx, const qmb_group *g, int n_up, int irrep, int layout, qmb_basis **out) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(out, "qmb_basis_build: out is NULL");
        *out = nullptr;
        QMB_REQUIRE(ctx && g, "qmb_basis_build: NULL context or group");
        QMB_REQUIRE(n_up >= 0 && n_up <= g->n_sites, "qmb_basis_build: n_up out of range");
        QMB_REQUIRE(n_up <= 32, "qmb_basis_build: at most 32 up spins (use the flipped sector)");
        QMB_REQUIRE(irrep >= 0 && irrep < g->n_irreps, "qmb_basis_build: irrep out of range");
        QMB_REQUIRE(g->irrep_ok[irrep], "qmb_basis_build: character row " + std::to_string(irrep) +
                                            " is not a one-dimensional representation of the group");
        QMB_REQUIRE(layout == QMB_LAYOUT_REFERENCE || layout == QMB_LAYOUT_COMPACT, "qmb_basis_build: bad layout");
        for (int e = 0; e < g->G; ++e)
            QMB_REQUIRE(!g->flip[e] || 2 * n_up == g->n_sites,
                        "qmb_basis_build: spin flip needs n_up == n_sites/2");
        const uint64_t dim = binom_host(g->n_sites, n_up);
        QMB_REQUIRE(dim > 0 && dim <= (1ULL << 38), "qmb_basis_build: sector too large for the rank dictionary");
        QMB_HIP(hipSetDevice(ctx->device));
        auto b = std::make_unique<qmb_basis>();
        b->ctx = ctx;
        b->group = g;
        b->n_sites = g->n_sites;
        b->n_up = n_up;
        b->irrep = irrep;
        b->layout = layout;
        b->sector_dim = (int64_t)dim;
        b->n_words = (int64_t)((dim + 63) / 64);
        const double *ch = &g->chars[(size_t)irrep * g->G * 2];
        b->all_rows = true;
        for (int e = 0; e < g->G; ++e) {
            const bool one = std::fabs(ch[2 * e] - 1.0) < 1e-9 && std::fabs(ch[2 * e + 1]) < 1e-9;
            b->all_rows = b->all_rows && one;
            if (one) {
                if (e < 64)
                    b->triv_lo |= 1ULL << e;
                else
                    b->triv_hi |= 1ULL << (e - 64);
            }
        }
        upload_group(b.get(), g);
        const GroupView gv = b->view();
        b->dict.alloc((size_t)b->n_words * 2 * sizeof(uint64_t));
        const unsigned blocks = (unsigned)((b->n_words + kBasisThreads - 1) / kBasisThreads);
        const bool narrow = g->n_sites <= 32;
        hipLaunchKernelGGL(narrow ? basis_mark<uint32_t> : basis_mark<uint64_t>, dim3(blocks), dim3(kBasisThreads), 0, ctx->stream, gv,
                           ctx->binom.as<uint64_t>(), n_up, b->sector_dim, b->n_words,
                           layout == QMB_LAYOUT_COMPACT ? 1 : 0, b->triv_lo, b->triv_hi, b->dict.as<uint64_t>());
        QMB_HIP(hipGetLastError());
        const int64_t rows = scan_popcount_dict(ctx, b->dict.as<uint64_t>(), b->n_words);
        QMB_REQUIRE(rows < (int64_t)INT32_MAX, "qmb_basis_build: more than 2^31 rows (int32 columns)");
        b->n_rows = rows;
        b->reps.alloc((size_t)std::max<int64_t>(rows, 1) * sizeof(uint64_t));
        b->stab.alloc((size_t)std::max<int64_t>(rows, 1));
        b->nullf.alloc((size_t)std::max<int64_t>(rows, 1));
        if (rows > 0) {
            const unsigned rblocks = (unsigned)((rows + kBasisThreads - 1) / kBasisThreads);
            hipLaunchKernelGGL(narrow ? basis_compact<uint32_t> : basis_compact<uint64_t>, dim3(rblocks), dim3(kBasisThreads), 0, ctx->stream, gv,
                               ctx->binom.as<uint64_t>(), n_up, rows, b->n_words, b->triv_lo, b->triv_hi,
                               b->dict.as<uint64_t>(), b->reps.as<uint64_t>(),
                               b->stab.as<uint8_t>(), b->nullf.as<uint8_t>());
            QMB_HIP(hipGetLastError());
        }
        // the compact layout returns with the compaction still queued (the ELL
        // build and every reader order after it on the stream)
        if (layout == QMB_LAYOUT_REFERENCE && rows > 0) {
            QMB_HIP(hipStreamSynchronize(ctx->stream));
            std::vector<uint8_t> nf(rows);
            QMB_HIP(hipMemcpy(nf.data(), b->nullf.p, rows, hipMemcpyDeviceToHost));
            int64_t nn = 0;
            for (auto v : nf) nn += v;
            b->n_null = nn;
        }
        *out = b.release();
    });
}

qmb_status qmb_basis_destroy(qmb_basis *b) {
    if (b && b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    qmb::release(b);
    return QMB_OK;
}

qmb_status qmb_basis_get_info(const qmb_basis *b, qmb_basis_info *info) {
    return guarded([&] {
        QMB_REQUIRE(b && info, "qmb_basis_get_info: NULL argument");
        info->sector_dim = b->sector_dim;
        info->n_rows = b->n_rows;
        info->n_null = b->n_null;
        info->n_sites = b->n_sites;
        info->n_up = b->n_up;
        info->irrep = b->irrep;
        info->layout = b->layout;
        info->device_bytes = (int64_t)(b->dict.bytes + b->reps.bytes + b->stab.bytes +
                                       b->nullf.bytes);
    });
}

qmb_status qmb_basis_export(const qmb_basis *b, uint64_t *reps, uint16_t *stab, uint8_t *is_null) {
    return guarded([&] {
        QMB_REQUIRE(b, "qmb_basis_export: NULL basis");
        QMB_HIP(hipSetDevice(b->ctx->device));
        QMB_HIP(hipStreamSynchronize(b->ctx->stream));
        const int64_t n = b->n_rows;
        if (n == 0) return;
        if (reps) QMB_HIP(hipMemcpy(reps, b->reps.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
        if (stab) {
            std::vector<uint8_t> s(n);
            QMB_HIP(hipMemcpy(s.data(), b->stab.p, n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) stab[i] = s[i];
        }
        if (is_null) QMB_HIP(hipMemcpy(is_null, b->nullf.p, n, hipMemcpyDeviceToHost));
    });
}

qmb_status qmb_basis_lookup(const qmb_basis *b, int64_t n, const uint64_t *states, int64_t *row, int32_t *elem) {
    return guarded([&] {
        PoolScope pool_scope(b && b->ctx ? &b->ctx->pool : nullptr);
        QMB_REQUIRE(b && states && row && elem, "qmb_basis_lookup: NULL argument");
        QMB_REQUIRE(n >= 0, "qmb_basis_lookup: negative count");
        if (n == 0) return;
        const uint64_t full = b->n_sites == 64 ? ~0ULL : ((1ULL << b->n_sites) - 1ULL);
        for (int64_t i = 0; i < n; ++i)
            QMB_REQUIRE((states[i] & ~full) == 0 && __builtin_popcountll(states[i]) == b->n_up,
                        "qmb_basis_lookup: state outside the sector");
        qmb_ctx *ctx = b->ctx;
        QMB_HIP(hipSetDevice(ctx->device));
        DevBuf ds(n * sizeof(uint64_t)), dr(n * sizeof(int64_t)), de(n * sizeof(int32_t));
        QMB_HIP(hipMemcpyAsync(ds.p, states, n * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
        const unsigned blocks = (unsigned)((n + kBasisThreads - 1) / kBasisThreads);
        hipLaunchKernelGGL(b->n_sites <= 32 ? basis_lookup<uint32_t> : basis_lookup<uint64_t>, dim3(blocks), dim3(kBasisThreads), 0, ctx->stream, b->view(),
                           ctx->binom.as<uint64_t>(), n, ds.as<uint64_t>(), b->dict.as<uint64_t>(),
                           dr.as<int64_t>(), de.as<int32_t>());
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipMemcpyAsync(row, dr.p, n * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipMemcpyAsync(elem, de.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
    });
}

}  // extern "C"
