// Device helpers shared by the basis and ELL kernels: group action, colex
// rank/unrank, next combination, orbit minimum and row lookup.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"

namespace qmb {

// Group tables staged in LDS by every block that applies group elements.
struct GroupLds {
    uint64_t keep[kMaxGroup];
    uint8_t small[kMaxGroup];
    uint8_t big[kMaxGroup];
    uint8_t flip[kMaxGroup];
    uint8_t inv[kMaxGroup];
};

// Binomial table C(p, j), p < 64, j <= 32, staged in LDS.
constexpr int kBinJ = 33;

__device__ inline void stage_group(GroupLds &lds, uint8_t *perm_lds, const GroupView &g) {
    for (int i = threadIdx.x; i < g.G; i += blockDim.x) {
        lds.keep[i] = g.keep[i];
        lds.small[i] = g.small[i];
        lds.big[i] = g.big[i];
        lds.flip[i] = g.flip[i];
        lds.inv[i] = g.inv[i];
    }
    if (!g.torus)
        for (int i = threadIdx.x; i < g.G * g.n_sites; i += blockDim.x) perm_lds[i] = g.perm[i];
}

__device__ inline void stage_binom(uint64_t *bin_lds, const uint64_t *bin) {
    for (int i = threadIdx.x; i < 64 * kBinJ; i += blockDim.x) bin_lds[i] = bin[i];
}

// Image of state x under element g: the reference's bit move s -> group_el[g][s]
// (Q/exact_diagonalization/utils.cu:52-59), plus the spin flip of this build.
__device__ inline uint64_t apply_elem(uint64_t x, int g, const GroupView &gv, const GroupLds &lds,
                                      const uint8_t *perm_lds) {
    uint64_t y;
    if (gv.torus) {
        y = x;
        const unsigned s = lds.small[g];
        if (s) {
            const uint64_t k = lds.keep[g];
            y = ((y & k) << s) | ((y & ~k & gv.full) >> (gv.block - s));
        }
        const unsigned b = lds.big[g];
        if (b) y = ((y << b) | (y >> (gv.n_sites - b))) & gv.full;
    } else {
        y = 0;
        const uint8_t *p = perm_lds + g * gv.n_sites;
        uint64_t t = x;
        while (t) {
            const int bit = __builtin_ctzll(t);
            y |= 1ULL << p[bit];
            t &= t - 1;
        }
    }
    if (lds.flip[g]) y ^= gv.full;
    return y;
}

// Colex rank: sum over set bits p_0 < p_1 < ... of C(p_j, j+1); equals the
// reference's GetIndexInHilbertSubspace (utils.cu:9-36).
__device__ inline int64_t colex_rank(uint64_t x, const uint64_t *bin_lds) {
    int64_t r = 0;
    int j = 1;
    while (x) {
        const int p = __builtin_ctzll(x);
        r += (int64_t)bin_lds[p * kBinJ + j];
        ++j;
        x &= x - 1;
    }
    return r;
}

// Inverse of colex_rank for popcount k (k <= 32).
__device__ inline uint64_t colex_unrank(int64_t rank, int k, int n_sites, const uint64_t *bin_lds) {
    uint64_t x = 0;
    uint64_t r = (uint64_t)rank;
    int p = n_sites - 1;
    for (int j = k; j >= 1; --j) {
        while (p >= j && bin_lds[p * kBinJ + j] > r) --p;
        if (p < j - 1) p = j - 1;
        x |= 1ULL << p;
        r -= bin_lds[p * kBinJ + j];
        --p;
    }
    return x;
}

// Next larger integer with the same popcount (ascending = colex order).
__device__ inline uint64_t next_comb(uint64_t x) {
    const uint64_t c = x & (~x + 1ULL);
    const uint64_t r = x + c;
    return (((r ^ x) >> 2) >> __builtin_ctzll(x)) | r;
}

// Representative test. Returns true when x is the smallest integer of its
// orbit; on success stab_lo/hi hold the stabiliser (elements with image x).
__device__ inline bool is_representative(uint64_t x, const GroupView &gv, const GroupLds &lds,
                                         const uint8_t *perm_lds, uint64_t &stab_lo,
                                         uint64_t &stab_hi) {
    stab_lo = 0;
    stab_hi = 0;
    for (int g = 0; g < gv.G; ++g) {
        const uint64_t y = apply_elem(x, g, gv, lds, perm_lds);
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
// deterministic).
__device__ inline uint64_t orbit_min(uint64_t x, const GroupView &gv, const GroupLds &lds,
                                     const uint8_t *perm_lds, int &gm) {
    uint64_t best = ~0ULL;
    int bi = kMaxGroup;
    for (int g = 0; g < gv.G; ++g) {
        const uint64_t y = apply_elem(x, g, gv, lds, perm_lds);
        const int ig = lds.inv[g];
        if (y < best) {
            best = y;
            bi = ig;
        } else if (y == best && ig < bi) {
            bi = ig;
        }
    }
    gm = bi;
    return best;
}

// Row of a representative through the basis bitmap (rank -> word -> prefix +
// popcount). Returns -1 when the representative is not a row.
__device__ inline int64_t row_of_rep(uint64_t rep, const uint64_t *bin_lds, const uint64_t *bitmap,
                                     const uint32_t *prefix) {
    const int64_t rk = colex_rank(rep, bin_lds);
    const uint64_t w = bitmap[rk >> 6];
    const int b = (int)(rk & 63);
    if (!((w >> b) & 1ULL)) return -1;
    const uint64_t below = b ? (w & ((1ULL << b) - 1ULL)) : 0ULL;
    return (int64_t)prefix[rk >> 6] + __popcll(below);
}

}  // namespace qmb
