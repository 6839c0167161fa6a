// Host check of the bit-sliced ring window scan (ring_windows in
// qmbcuda_amd/csrc/device_common.hpp) against a plain scan of the 8-bit windows
// of every rotation (and complement): same minimum, same candidate positions,
// on random states of 8 to 56 sites with and without the spin flip.
#include <cstdio>
#include <random>

#include "device_common.hpp"

static unsigned scan(uint64_t y, int N, bool hf, uint64_t &c0, uint64_t &c1) {
    const uint64_t yy = y | (y << N);
    unsigned tmin = 0x1FFu;
    c0 = c1 = 0;
    for (int p = 0; p < N; ++p) {
        const unsigned w = (unsigned)(yy >> p) & 0xFFu;
        if (w < tmin) {
            tmin = w;
            c0 = c1 = 0;
        }
        if (w == tmin) c0 |= 1ULL << p;
        if (hf) {
            const unsigned wf = w ^ 0xFFu;
            if (wf < tmin) {
                tmin = wf;
                c0 = c1 = 0;
            }
            if (wf == tmin) c1 |= 1ULL << p;
        }
    }
    return tmin;
}

int main() {
    std::mt19937_64 rng(11);
    long bad = 0, n = 0;
    for (int it = 0; it < 400000; ++it) {
        const int N = 8 + (int)(rng() % 49);
        const bool hf = rng() & 1;
        uint64_t y = rng() & ((N == 64) ? ~0ULL : ((1ULL << N) - 1ULL));
        if (it % 5 == 0) y &= rng();  // sparser states
        uint64_t a0, a1, b0, b1;
        const unsigned t1 = scan(y, N, hf, a0, a1);
        unsigned t2;
        if (N <= 32)
            t2 = qmb::ring_windows<uint32_t>((uint32_t)y, N, hf, b0, b1);
        else
            t2 = qmb::ring_windows<uint64_t>(y, N, hf, b0, b1);
        ++n;
        if (t1 != t2 || a0 != b0 || a1 != b1) ++bad;
    }
    std::printf("states %ld mismatches %ld\n", n, bad);
    return bad == 0 ? 0 : 1;
}
