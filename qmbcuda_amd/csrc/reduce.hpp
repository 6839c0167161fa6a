// Deterministic block reductions shared by the Lanczos kernels: wave shuffles,
// a fixed-order LDS combine and a fixed-order sum of per-block partial slots
// (run by a separate one-block kernel; no floating-point atomics, no
// device-scope fences inside the streaming kernels). Also the counter-based
// generator of the seeded start vectors.
#pragma once

#include <hip/hip_runtime.h>

namespace qmb {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum over the block, valid in thread 0.
__device__ inline double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) lds[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWaves; ++i) t += lds[i];
    __syncthreads();
    return t;
}

// Deterministic sum of partials[0..n) by one block (fixed order): thread t adds
// partials t, t + blockDim, t + 2 blockDim, ... in that order; eight loads are
// issued before they are added, so the latency of the L2 reads overlaps.
__device__ inline double sum_partials(const double *partials, int n, double *lds) {
    double s = 0.0;
    const int stride = blockDim.x;
    int i = threadIdx.x;
    for (; i + 7 * stride < n; i += 8 * stride) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partials[i + u * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; i < n; i += stride) s += partials[i];
    return block_sum(s, lds);
}

// Counter-based SplitMix64 (Steele, Lea, Flood 2014) -> uniform (0, 1]:
// element c of the stream of `seed`, independent of the launch shape.
__device__ inline double splitmix_uniform(uint64_t seed, uint64_t c) {
    uint64_t z = seed + (c + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)((z >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
}

}  // namespace qmb
