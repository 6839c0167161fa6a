// Host check of the nibble-table colex rank (colex_rank_nib and
// rank_nibble_table in qmbcuda_amd/csrc/device_common.hpp) against the
// definition of the colex rank, the sum over the set bits p_0 < p_1 < ... of
// C(p_j, j + 1) (the reference's GetIndexInHilbertSubspace,
// Q/exact_diagonalization/utils.cu:9-36): every state of up to 16 sites, and
// random states of larger lattices up to 64 sites.
#include <cstdio>
#include <random>

#include "device_common.hpp"

using namespace qmb;

static uint64_t C[65][65];

static uint64_t rank_def(uint64_t x) {
    uint64_t r = 0;
    int j = 1;
    for (int p = 0; p < 64; ++p)
        if ((x >> p) & 1ULL) r += C[p][j++];
    return r;
}

// a random state of n sites with k set bits
static uint64_t random_state(std::mt19937_64 &rng, int n, int k) {
    uint64_t x = 0;
    int have = 0;
    while (have < k) {
        const int p = (int)(rng() % (uint64_t)n);
        if (!((x >> p) & 1ULL)) {
            x |= 1ULL << p;
            ++have;
        }
    }
    return x;
}

int main() {
    for (int n = 0; n <= 64; ++n) {
        C[n][0] = 1;
        for (int k = 1; k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k < n ? C[n - 1][k] : 0);
    }
    long checked = 0, mismatches = 0;
    for (int n = 1; n <= 16; ++n)
        for (int k = 0; k <= n; ++k) {
            int nn = 0, nc = 0;
            const std::vector<uint64_t> t = rank_nibble_table(n, k, nn, nc);
            for (uint64_t x = 0; x < (1ULL << n); ++x) {
                if (__builtin_popcountll(x) != k) continue;
                ++checked;
                if ((uint64_t)colex_rank_nib<uint64_t>(x, t.data(), nn, nc) != rank_def(x)) ++mismatches;
            }
        }
    std::mt19937_64 rng(11);
    const int cases[][2] = {{28, 14}, {30, 15}, {33, 16}, {36, 18}, {36, 12}, {40, 20}, {48, 24}, {64, 32}};
    for (const auto &cs : cases) {
        int nn = 0, nc = 0;
        const std::vector<uint64_t> t = rank_nibble_table(cs[0], cs[1], nn, nc);
        for (int i = 0; i < 200000; ++i) {
            const uint64_t x = random_state(rng, cs[0], cs[1]);
            ++checked;
            if ((uint64_t)colex_rank_nib<uint64_t>(x, t.data(), nn, nc) != rank_def(x)) ++mismatches;
            if (cs[0] <= 32 && (uint64_t)colex_rank_nib<uint32_t>((uint32_t)x, t.data(), nn, nc) != rank_def(x))
                ++mismatches;
        }
    }
    std::printf("checked %ld states, mismatches %ld\n", checked, mismatches);
    return mismatches ? 1 : 0;
}
