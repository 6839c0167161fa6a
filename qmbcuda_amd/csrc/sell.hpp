// SELL-64 row products shared by the single and the batched SpMV kernels.
//
// One lane owns one row of a 64-row chunk; the chunk is stored slot-major, so
// slot k of the 64 rows is one coalesced 256 B (col) + 512 B / 1 KB (val)
// read. Slots are taken four at a time: the four column and value loads are
// issued together, then the four x gathers, then the four products are added
// in slot order, so the sum has the same rounding sequence as a plain
// slot-by-slot loop (and as the FP64 oracle's row sum).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qmb {

// Matrix-stream load: nontemporal (streamed past the caches) when NT, else
// plain (the lines may stay in L2 and the MALL for the next SpMV).
template <bool NT, class T>
__device__ inline T stream_load(const T *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// sum over slots k < len of val[p0 + 64k] * x[col[p0 + 64k]] (L = chunk slots)
template <bool NT>
__device__ inline double sell_row_real(const int32_t *__restrict__ col, const double *__restrict__ val,
                                       const double *__restrict__ x, int64_t p0, int L, int len) {
    double acc = 0.0;
    // next group's columns and values loaded before this group's gathers
    const int n4 = L / 4;
    int32_t jn[4] = {0, 0, 0, 0};
    double vn[4] = {0.0, 0.0, 0.0, 0.0};
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k0 + u < len) {
                const int64_t p = p0 + (int64_t)(k0 + u) * 64;
                jn[u] = stream_load<NT>(col + p);
                vn[u] = stream_load<NT>(val + p);
            }
        }
    };
    if (n4 > 0) load(0);
    int k = 0;
    for (int q = 0; q < n4; ++q, k += 4) {
        int32_t j[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            j[u] = jn[u];
            v[u] = vn[u];
        }
        if (q + 1 < n4) load(k + 4);
        double g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < len) g[u] = x[j[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < len) acc += v[u] * g[u];
    }
    for (; k < L; ++k) {
        if (k < len) {
            const int64_t p = p0 + (int64_t)k * 64;
            const int32_t jj = stream_load<NT>(col + p);
            const double vv = stream_load<NT>(val + p);
            acc += vv * x[jj];
        }
    }
    return acc;
}

// Packed real rows: slot 0 is the diagonal d (column = the row r itself); slot
// k >= 1 holds w = col | (index << cb) with the value vt[index]. Same products in
// the same order as sell_row_real on the unpacked matrix.
// Slots are taken in groups of U from slot 1 (8 when the matrix is streamed
// from HBM: more loads in flight; 4 when it is cached); the next group's words
// are loaded before this group's gathers, so the two latencies overlap.
template <bool NT, int U = (NT ? 8 : 4)>
__device__ inline double sell_row_packed(const uint32_t *__restrict__ w, const double *vt,
                                         const double *__restrict__ x, int64_t p0, int L, int len, int cb,
                                         double d, double xr) {
    const uint32_t mask = (1u << cb) - 1u;
    double acc = 0.0;
    if (len > 0) acc += d * xr;
    const int n4 = (L - 1) / U;
    uint32_t jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = 0u;
    if (n4 > 0) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (1 + u < len) jn[u] = stream_load<NT>(w + p0 + (int64_t)(1 + u) * 64);
    }
    int k = 1;
    for (int q = 0; q < n4; ++q, k += U) {
        uint32_t j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
        if (q + 1 < n4) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + U + u < len) jn[u] = stream_load<NT>(w + p0 + (int64_t)(k + U + u) * 64);
        }
        double g[U], v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < len) {
                g[u] = x[j[u] & mask];
                v[u] = vt[j[u] >> cb];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k + u < len) acc += v[u] * g[u];
    }
    for (; k < L; ++k) {
        if (k < len) {
            const uint32_t jj = stream_load<NT>(w + p0 + (int64_t)k * 64);
            acc += vt[jj >> cb] * x[jj & mask];
        }
    }
    return acc;
}

// complex rows: val and x are interleaved (re, im) pairs
template <bool NT>
__device__ inline double2 sell_row_cplx(const int32_t *__restrict__ col, const double *__restrict__ val,
                                        const double *__restrict__ x, int64_t p0, int L, int len) {
    const double2 *val2 = reinterpret_cast<const double2 *>(val);
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    double ar = 0.0, ai = 0.0;
    int k = 0;
    for (; k + 4 <= L; k += 4) {
        int32_t j[4];
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k + u < len) {
                const int64_t p = p0 + (int64_t)(k + u) * 64;
                j[u] = stream_load<NT>(col + p);
                v[u] = val2[p];
            }
        }
        double2 g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k + u < len) g[u] = x2[j[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k + u < len) {
                ar += v[u].x * g[u].x - v[u].y * g[u].y;
                ai += v[u].x * g[u].y + v[u].y * g[u].x;
            }
        }
    }
    for (; k < L; ++k) {
        if (k < len) {
            const int64_t p = p0 + (int64_t)k * 64;
            const int32_t jj = stream_load<NT>(col + p);
            const double2 vv = val2[p];
            const double2 gg = x2[jj];
            ar += vv.x * gg.x - vv.y * gg.y;
            ai += vv.x * gg.y + vv.y * gg.x;
        }
    }
    return make_double2(ar, ai);
}

// Packed complex rows: slot 0 is the diagonal d; slot k >= 1 holds
// w = col | (re index << cb) | (im index << (cb + rb)) with the value
// (vre[re index], vim[im index]). Same products in the same order as
// sell_row_cplx on the unpacked matrix; groups of U slots as in sell_row_packed.
template <bool NT, int U = 4>
__device__ inline double2 sell_row_cplx_packed(const uint32_t *__restrict__ w, const double *vre, const double *vim,
                                               const double *__restrict__ x, int64_t p0, int L, int len, int cb,
                                               int rb, double2 d, double2 xr) {
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    const uint32_t mask = (1u << cb) - 1u, rmask = (1u << rb) - 1u;
    const int sh2 = cb + rb;
    double ar = 0.0, ai = 0.0;
    if (len > 0) {
        ar += d.x * xr.x - d.y * xr.y;
        ai += d.x * xr.y + d.y * xr.x;
    }
    // next group's words loaded before this group's gathers (as in sell_row_packed)
    const int n4 = (L - 1) / U;
    uint32_t jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = 0u;
    if (n4 > 0) {
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (1 + u < len) jn[u] = stream_load<NT>(w + p0 + (int64_t)(1 + u) * 64);
    }
    int k = 1;
    for (int q = 0; q < n4; ++q, k += U) {
        uint32_t j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
        if (q + 1 < n4) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k + U + u < len) jn[u] = stream_load<NT>(w + p0 + (int64_t)(k + U + u) * 64);
        }
        double2 g[U], v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < len) {
                g[u] = x2[j[u] & mask];
                v[u].x = vre[(j[u] >> cb) & rmask];
                v[u].y = vim[(uint32_t)((uint64_t)j[u] >> sh2)];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < len) {
                ar += v[u].x * g[u].x - v[u].y * g[u].y;
                ai += v[u].x * g[u].y + v[u].y * g[u].x;
            }
        }
    }
    for (; k < L; ++k) {
        if (k < len) {
            const uint32_t jj = stream_load<NT>(w + p0 + (int64_t)k * 64);
            const double vr = vre[(jj >> cb) & rmask], vi = vim[(uint32_t)((uint64_t)jj >> sh2)];
            const double2 gg = x2[jj & mask];
            ar += vr * gg.x - vi * gg.y;
            ai += vr * gg.y + vi * gg.x;
        }
    }
    return make_double2(ar, ai);
}

}  // namespace qmb
