// Device helpers shared by the basis and ELL kernels: group action, colex
// rank/unrank, next combination, orbit minimum and row lookup.
//
// S is the state word: uint32_t when the lattice has at most 32 sites (half the
// VALU work of 64-bit words), uint64_t otherwise. The group loops are uniform
// across a wavefront, so the per-element tables are read straight from global
// memory (scalar loads); only lane-divergent tables (binomials, generic
// permutations) are staged in LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace qmb {

// Binomial table C(p, j), p < 64, j <= 32, staged in LDS.
constexpr int kBinJ = 33;

__device__ inline int ctz_s(uint32_t x) { return __builtin_ctz(x); }
__device__ inline int ctz_s(uint64_t x) { return __builtin_ctzll(x); }
__device__ inline int popc_s(uint32_t x) { return __popc(x); }
__device__ inline int popc_s(uint64_t x) { return __popcll(x); }

__device__ inline void stage_perm(uint8_t *perm_lds, const GroupView &g) {
    if (!g.torus)
        for (int i = threadIdx.x; i < g.G * g.n_sites; i += blockDim.x) perm_lds[i] = g.perm[i];
}

__device__ inline void stage_binom(uint64_t *bin_lds, const uint64_t *bin) {
    for (int i = threadIdx.x; i < 64 * kBinJ; i += blockDim.x) bin_lds[i] = bin[i];
}

template <class S>
__device__ inline S full_mask(const GroupView &gv) {
    return (S)gv.full;
}

// Image of state x under element g: the reference's bit move s -> group_el[g][s]
// (Q/exact_diagonalization/utils.cu:52-59), plus the spin flip of this build.
template <class S>
__device__ inline S apply_elem(S x, int g, const GroupView &gv, const uint8_t *perm_lds) {
    const S full = (S)gv.full;
    S y;
    if (gv.torus) {
        y = x;
        const unsigned s = gv.small[g];
        if (s) {
            const S k = (S)gv.keep[g];
            y = (S)(((y & k) << s) | ((y & (S)~k & full) >> (gv.block - s)));
        }
        const unsigned b = gv.big[g];
        if (b) y = (S)(((y << b) | (y >> (gv.n_sites - b))) & full);
    } else {
        y = 0;
        const uint8_t *p = perm_lds + g * gv.n_sites;
        S t = x;
        while (t) {
            const int bit = ctz_s(t);
            y |= (S)1 << p[bit];
            t &= t - 1;
        }
    }
    if (gv.flip[g]) y ^= full;
    return y;
}

// LDS copy of the per-element tables for lane-divergent element indices, and
// (when the group has them) the torus orbit-minimum tables.
struct GroupLds {
    uint64_t keep[kMaxGroup];
    uint8_t small[kMaxGroup];
    uint8_t big[kMaxGroup];
    uint8_t flip[kMaxGroup];
    uint8_t inv[kMaxGroup];
    uint8_t lut[kLutBytes];
    uint8_t cos[kCosBytes];
};

__device__ inline void stage_group(GroupLds &t, const GroupView &g) {
    for (int i = threadIdx.x; i < g.G; i += blockDim.x) {
        t.keep[i] = g.keep[i];
        t.small[i] = g.small[i];
        t.big[i] = g.big[i];
        t.flip[i] = g.flip[i];
        t.inv[i] = g.inv[i];
    }
    if (g.lut == 1 || g.lut == 2)
        for (int i = threadIdx.x; i < kLutBytes; i += blockDim.x) t.lut[i] = g.lut_tab[i];
    if (g.lut == 3)
        for (int i = threadIdx.x; i < kCosBytes; i += blockDim.x) t.cos[i] = g.cos_tab[i];
}

// apply_elem with the element index varying across lanes (tables from LDS).
template <class S>
__device__ inline S apply_elem_div(S x, int g, const GroupView &gv, const GroupLds &t, const uint8_t *perm_lds) {
    const S full = (S)gv.full;
    S y;
    if (gv.torus) {
        y = x;
        const unsigned s = t.small[g];
        if (s) {
            const S k = (S)t.keep[g];
            y = (S)(((y & k) << s) | ((y & (S)~k & full) >> (gv.block - s)));
        }
        const unsigned b = t.big[g];
        if (b) y = (S)(((y << b) | (y >> (gv.n_sites - b))) & full);
    } else {
        y = 0;
        const uint8_t *p = perm_lds + g * gv.n_sites;
        S u = x;
        while (u) {
            const int bit = ctz_s(u);
            y |= (S)1 << p[bit];
            u &= u - 1;
        }
    }
    if (t.flip[g]) y ^= full;
    return y;
}

// Colex rank: sum over set bits p_0 < p_1 < ... of C(p_j, j+1); equals the
// reference's GetIndexInHilbertSubspace (utils.cu:9-36).
template <class S>
__device__ inline int64_t colex_rank(S x, const uint64_t *bin_lds) {
    int64_t r = 0;
    int j = 1;
    while (x) {
        const int p = ctz_s(x);
        r += (int64_t)bin_lds[p * kBinJ + j];
        ++j;
        x &= x - 1;
    }
    return r;
}

// Inverse of colex_rank for popcount k (k <= 32).
template <class S>
__device__ inline S colex_unrank(int64_t rank, int k, int n_sites, const uint64_t *bin_lds) {
    S x = 0;
    uint64_t r = (uint64_t)rank;
    int p = n_sites - 1;
    for (int j = k; j >= 1; --j) {
        while (p >= j && bin_lds[p * kBinJ + j] > r) --p;
        if (p < j - 1) p = j - 1;
        x |= (S)1 << p;
        r -= bin_lds[p * kBinJ + j];
        --p;
    }
    return x;
}

// Next larger integer with the same popcount (ascending = colex order).
template <class S>
__device__ inline S next_comb(S x) {
    const S c = x & (S)(~x + (S)1);
    const S r = x + c;
    return (S)((((r ^ x) >> 2) >> ctz_s(x)) | r);
}

// Image of x under coset representative c_i (i >= 1) of a lut-3 group: one bit
// reversal when c_i is the site reversal, else the generic permutation.
template <class S>
__device__ inline S reverse_sites(S x, int n);
template <>
__device__ inline uint32_t reverse_sites<uint32_t>(uint32_t x, int n) {
    return __builtin_bitreverse32(x) >> (32 - n);
}
template <>
__device__ inline uint64_t reverse_sites<uint64_t>(uint64_t x, int n) {
    return __builtin_bitreverse64(x) >> (64 - n);
}
template <class S>
__device__ inline S coset_image(S x, int ci, const GroupView &gv, const GroupLds &t, const uint8_t *perm_lds) {
    if ((gv.cos_rev >> ci) & 1) return reverse_sites<S>(x, gv.n_sites);
    return apply_elem_div<S>(x, t.cos[ci * 128], gv, t, perm_lds);
}

// Left rotation of an n-site ring state by s (0 <= s < n).
template <class S>
__device__ inline S rotl_ring(S y, int s, int n, S full) {
    return s ? (S)(((y << s) | (y >> (n - s))) & full) : y;
}

// Stabiliser of x for a group of ring rotations times cosets (gv.lut == 3):
// the images of (rotation s, flip f) o c_i are the rotations of c_i x.
template <class S>
__device__ inline void stabiliser_coset(S x, const GroupView &gv, const uint8_t *perm_lds, uint64_t &stab_lo,
                                        uint64_t &stab_hi) {
    const int N = gv.n_sites;
    const S full = (S)gv.full;
    stab_lo = 0;
    stab_hi = 0;
    for (int ci = 0; ci < gv.n_cos; ++ci) {
        const uint8_t *ct = gv.cos_tab + ci * 128;
        S y = !ci                        ? x
              : ((gv.cos_rev >> ci) & 1) ? reverse_sites<S>(x, N)
                                         : apply_elem<S>(x, ct[0], gv, perm_lds);
        for (int s = 0; s < N; ++s) {
            if (s) y = rotl_ring<S>(y, 1, N, full);
            for (int f = 0; f < (gv.lut_flip ? 2 : 1); ++f) {
                if ((f ? (S)(y ^ full) : y) != x) continue;
                const int g = ct[f * 64 + s];
                if (g < 64)
                    stab_lo |= 1ULL << g;
                else
                    stab_hi |= 1ULL << (g - 64);
            }
        }
    }
}

// Representative test. Returns true when x is the smallest integer of its
// orbit; on success stab_lo/hi hold the stabiliser (elements with image x).
template <class S>
__device__ inline bool is_representative(S x, const GroupView &gv, const uint8_t *perm_lds, uint64_t &stab_lo,
                                         uint64_t &stab_hi) {
    if (gv.lut == 3) {
        stabiliser_coset<S>(x, gv, perm_lds, stab_lo, stab_hi);
        return true;  // callers pass representatives only
    }
    if (gv.lut == 2) {
        // ring: the rotations one bit at a time (and their complements); the
        // element table is read only for images equal to x
        const int N = gv.n_sites;
        const S full = (S)gv.full;
        stab_lo = 0;
        stab_hi = 0;
        S y = x;
        for (int st = 0; st < N; ++st) {
            if (st) y = rotl_ring<S>(y, 1, N, full);
            for (int f = 0; f < (gv.lut_flip ? 2 : 1); ++f) {
                if ((f ? (S)(y ^ full) : y) != x) continue;
                const int g = gv.lut_tab[512 + f * 64 + st];
                if (g < 64)
                    stab_lo |= 1ULL << g;
                else
                    stab_hi |= 1ULL << (g - 64);
            }
        }
        return true;  // callers pass representatives only
    }
    stab_lo = 0;
    stab_hi = 0;
    for (int g = 0; g < gv.G; ++g) {
        const S y = apply_elem<S>(x, g, gv, perm_lds);
        if (y < x) return false;
        if (y == x) {
            if (g < 64)
                stab_lo |= 1ULL << g;
            else
                stab_hi |= 1ULL << (g - 64);
        }
    }
    return true;
}

// Orbit minimum of a state and the canonical element gm = the smallest g with
// image_g(min) == x (the reference's orbit_indices entry, utils.cu:66-80, made
// deterministic). mult = the number of elements mapping x to the minimum, i.e.
// the stabiliser size of the minimum (those elements form a coset of x's
// stabiliser, which has the same size as the minimum's).
template <class S>
__device__ inline S orbit_min(S x, const GroupView &gv, const uint8_t *perm_lds, int &gm, int &mult) {
    S best = x;
    int bi = gv.inv[0];
    mult = 1;
    for (int g = 1; g < gv.G; ++g) {
        const S y = apply_elem<S>(x, g, gv, perm_lds);
        const int ig = gv.inv[g];
        if (y < best) {
            best = y;
            bi = ig;
            mult = 1;
        } else if (y == best) {
            ++mult;
            if (ig < bi) bi = ig;
        }
    }
    gm = bi;
    return best;
}

// orbit_min for a group that is the whole torus product (gv.lut): an image's
// top block is source block j rotated by the in-block shift s (complemented
// under flip), placed on top by the block shift a = nb-1-j. Integer order is
// decided by the top block first, so the minimum image has the smallest top
// block tmin over all (j, s, f), and only the elements reaching tmin need full
// images. Same result and same gm as orbit_min.
template <class S>
__device__ inline S orbit_min_lut(S x, const GroupView &gv, const GroupLds &t, int &gm, int &mult) {
    const int B = gv.block, N = gv.n_sites, nb = N / B;
    const unsigned mB = (1u << B) - 1u;
    const S full = (S)gv.full;
    const int nf = gv.lut_flip ? 2 : 1;
    unsigned tmin = 0xFFu;
    for (int j = 0; j < nb; ++j) {
        const unsigned v = (unsigned)(x >> (j * B)) & mB;
        tmin = min(tmin, (unsigned)t.lut[v]);
        if (nf == 2) tmin = min(tmin, (unsigned)t.lut[v ^ mB]);
    }
    S best = 0;
    int bi = -1;
    mult = 0;
    for (int j = 0; j < nb; ++j) {
        const unsigned v0 = (unsigned)(x >> (j * B)) & mB;
        const int a = nb - 1 - j;
        for (int f = 0; f < nf; ++f) {
            const unsigned v = f ? (v0 ^ mB) : v0;
            if (t.lut[v] != tmin) continue;
            unsigned sm = t.lut[256 + v];
            const int big = a * B;
            while (sm) {
                const int sh = __builtin_ctz(sm);
                sm &= sm - 1;
                // the element (in-block shift sh, block shift a, flip f) applied
                // directly (as apply_elem): its tables are read only for a new
                // or tied minimum
                S y = x;
                if (sh) {
                    const S k = (S)((S)gv.rep_unit * (S)((1u << (B - sh)) - 1u));
                    y = (S)(((y & k) << sh) | ((y & (S)~k & full) >> (B - sh)));
                }
                if (big) y = (S)(((y << big) | (y >> (N - big))) & full);
                if (f) y ^= full;
                if (bi < 0 || y < best) {
                    best = y;
                    bi = t.inv[t.lut[512 + (f * 8 + a) * 8 + sh]];
                    mult = 1;
                } else if (y == best) {
                    ++mult;
                    const int ig = t.inv[t.lut[512 + (f * 8 + a) * 8 + sh]];
                    if (ig < bi) bi = ig;
                }
            }
        }
    }
    gm = bi;
    return best;
}

// Representative test for a full torus product (gv.lut == 1): x is the orbit
// minimum iff no image is smaller. Only the images whose top block reaches the
// smallest rotated block tmin can be <= x (the others have a larger top block),
// so those candidates decide it, and the ones equal to x are the stabiliser.
// The caller has checked that x's own top block is tmin.
template <class S>
__device__ inline bool rep_stab_lut(S x, const GroupView &gv, const GroupLds &t, uint64_t &stab_lo,
                                    uint64_t &stab_hi) {
    const int B = gv.block, nb = gv.n_sites / B;
    const unsigned mB = (1u << B) - 1u;
    const unsigned tmin = (unsigned)(x >> ((nb - 1) * B)) & mB;
    stab_lo = 0;
    stab_hi = 0;
    for (int j = 0; j < nb; ++j) {
        const unsigned v0 = (unsigned)(x >> (j * B)) & mB;
        const int a = nb - 1 - j;
        for (int f = 0; f < (gv.lut_flip ? 2 : 1); ++f) {
            const unsigned v = f ? (v0 ^ mB) : v0;
            if (t.lut[v] != tmin) continue;
            unsigned sm = t.lut[256 + v];
            while (sm) {
                const int sh = __builtin_ctz(sm);
                sm &= sm - 1;
                const int g = t.lut[512 + (f * 8 + a) * 8 + sh];
                const S y = apply_elem_div<S>(x, g, gv, t, nullptr);
                if (y < x) return false;
                if (y == x) {
                    if (g < 64)
                        stab_lo |= 1ULL << g;
                    else
                        stab_hi |= 1ULL << (g - 64);
                }
            }
        }
    }
    return true;
}

// orbit_min for a ring group (gv.lut == 2): ring_windows finds the smallest
// top byte over the rotations (and complements) and the lane's candidate
// positions; full images are taken of the candidates only. Same result and
// same gm as orbit_min.
template <class S>
__host__ __device__ inline unsigned ring_windows(S y, int N, bool hf, uint64_t &c0, uint64_t &c1);

template <class S>
__device__ inline S orbit_min_ring(S x, const GroupView &gv, const GroupLds &t, int &gm, int &mult) {
    const int N = gv.n_sites;
    const S full = (S)gv.full;
    const bool hf = gv.lut_flip != 0;
    uint64_t c0, c1;
    ring_windows<S>(x, N, hf, c0, c1);
    S best = 0;
    int bi = -1;
    mult = 0;
    for (int f = 0; f < (hf ? 2 : 1); ++f) {
        uint64_t c = f ? c1 : c0;
        while (c) {
            const int p = __builtin_ctzll(c);
            c &= c - 1;
            int s = N - 8 - p;
            if (s < 0) s += N;
            // rotation by s (complemented under f) in registers; the element
            // tables are read only for a new or tied minimum
            S y = rotl_ring<S>(x, s, N, full);
            if (f) y ^= full;
            if (bi < 0 || y < best) {
                best = y;
                bi = t.inv[t.lut[512 + f * 64 + s]];
                mult = 1;
            } else if (y == best) {
                ++mult;
                const int ig = t.inv[t.lut[512 + f * 64 + s]];
                if (ig < bi) bi = ig;
            }
        }
    }
    gm = bi;
    return best;
}

// Window scan of the ring orbit minimum: the smallest top byte over the
// rotations of y (and their complements when hf) and the window positions p
// that reach it (c0: plain, c1: complemented). The top 8 bits of rotation s
// are the circular window of y starting at site p = (N - 8 - s) mod N, whose
// bit i is bit p + i of the doubled word. All windows are compared at once,
// bit-sliced from the top bit down: at bit i the candidates whose window bit
// is 0 (complemented: 1) survive if there are any, else bit i of the minimum
// is 1 and every candidate stays.
template <class S>
__host__ __device__ inline unsigned ring_windows(S y, int N, bool hf, uint64_t &c0, uint64_t &c1) {
    const uint64_t yy = (uint64_t)y | ((uint64_t)y << N);  // 8 <= N <= 56: bits p + i < 64
    const uint64_t all = (1ull << N) - 1ull;
    uint64_t a = all, b = hf ? all : 0ull;
    unsigned tmin = 0;
#pragma unroll
    for (int i = 7; i >= 0; --i) {
        const uint64_t bi = yy >> i;
        const uint64_t z0 = a & ~bi, z1 = b & bi;
        const bool any = (z0 | z1) != 0ull;
        a = any ? z0 : a;
        b = any ? z1 : b;
        tmin |= any ? 0u : (1u << i);
    }
    c0 = a;
    c1 = b;
    return tmin;
}

// orbit_min for ring rotations times cosets (gv.lut == 3): the orbit is the
// union of the ring orbits of y_i = c_i x. Each coset's window scan gives its
// smallest top byte; only cosets reaching the smallest so far take full images
// of their candidate rotations. Same result and same gm as orbit_min.
template <class S>
__device__ inline S orbit_min_coset(S x, const GroupView &gv, const GroupLds &t, const uint8_t *perm_lds,
                                    int &gm, int &mult) {
    const int N = gv.n_sites;
    const S full = (S)gv.full;
    const bool hf = gv.lut_flip != 0;
    S best = 0;
    int bi = -1;
    mult = 0;
    unsigned tbest = 0x1FFu;
    for (int ci = 0; ci < gv.n_cos; ++ci) {
        const uint8_t *ct = t.cos + ci * 128;
        const S y = ci ? coset_image<S>(x, ci, gv, t, perm_lds) : x;
        uint64_t c0, c1;
        const unsigned tmin = ring_windows<S>(y, N, hf, c0, c1);
        if (tmin > tbest) continue;
        tbest = tmin;
        for (int f = 0; f < (hf ? 2 : 1); ++f) {
            uint64_t c = f ? c1 : c0;
            while (c) {
                const int p = __builtin_ctzll(c);
                c &= c - 1;
                int s = N - 8 - p;
                if (s < 0) s += N;
                const int g = ct[f * 64 + s];
                S z = rotl_ring<S>(y, s, N, full);
                if (f) z ^= full;
                const int ig = t.inv[g];
                if (bi < 0 || z < best) {
                    best = z;
                    bi = ig;
                    mult = 1;
                } else if (z == best) {
                    ++mult;
                    if (ig < bi) bi = ig;
                }
            }
        }
    }
    gm = bi;
    return best;
}

// Colex rank by nibbles: T[(k * n_c + c) * 16 + v] is the rank contribution of
// nibble k (sites 4k..4k+3) holding v when c set bits lie below it, i.e. the sum
// over v's set bits i (ascending, j of v's bits below i) of C(4k + i, c + j + 1).
// ceil(n_sites / 4) table reads instead of one per set bit; same value as
// colex_rank for every state with at most n_c - 1 set bits.
template <class S>
__host__ __device__ inline int64_t colex_rank_nib(S x, const uint64_t *tab, int n_nib, int n_c) {
    int64_t r = 0;
    int c = 0;
    for (int k = 0; k < n_nib; ++k) {
        const unsigned v = (unsigned)(x >> (4 * k)) & 15u;
        r += (int64_t)tab[(k * n_c + c) * 16 + (int)v];
        c += __builtin_popcount(v);
    }
    return r;
}

// Host: the nibble rank table for n_sites sites and states of n_up set bits
// (n_nib = ceil(n_sites / 4), n_c = n_up + 1); entries that no such state reads
// (more than n_up bits up to that nibble) hold the same formula.
inline std::vector<uint64_t> rank_nibble_table(int n_sites, int n_up, int &n_nib, int &n_c) {
    n_nib = (n_sites + 3) / 4;
    n_c = n_up + 1;
    const int P = 4 * n_nib + 1;
    std::vector<std::vector<uint64_t>> C(P, std::vector<uint64_t>(P + 8, 0));
    for (int n = 0; n < P; ++n) {
        C[n][0] = 1;
        for (int k = 1; k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k <= n - 1 ? C[n - 1][k] : 0);
    }
    auto binom = [&](int n, int k) -> uint64_t { return (k < 0 || k > n) ? 0 : C[n][k]; };
    std::vector<uint64_t> t((size_t)n_nib * n_c * 16, 0);
    for (int k = 0; k < n_nib; ++k)
        for (int c = 0; c < n_c; ++c)
            for (int v = 0; v < 16; ++v) {
                uint64_t r = 0;
                int j = 0;
                for (int i = 0; i < 4; ++i)
                    if ((v >> i) & 1) {
                        r += binom(4 * k + i, c + j + 1);
                        ++j;
                    }
                t[((size_t)k * n_c + c) * 16 + v] = r;
            }
    return t;
}

// Row of a representative through the rank dictionary (rank -> record -> rows
// before the word + popcount of the lower bits). Returns -1 when the
// representative is not a row.
__device__ inline int64_t row_of_rank(int64_t rk, const uint64_t *dict) {
    const ulonglong2 rec = reinterpret_cast<const ulonglong2 *>(dict)[rk >> 6];
    const uint64_t w = rec.x;
    const int b = (int)(rk & 63);
    if (!((w >> b) & 1ULL)) return -1;
    const uint64_t below = b ? (w & ((1ULL << b) - 1ULL)) : 0ULL;
    return (int64_t)rec.y + __popcll(below);
}

template <class S>
__device__ inline int64_t row_of_rep(S rep, const uint64_t *bin_lds, const uint64_t *dict) {
    return row_of_rank(colex_rank<S>(rep, bin_lds), dict);
}

}  // namespace qmb
