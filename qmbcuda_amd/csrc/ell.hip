// On-device ELL Hamiltonian assembly.
//
// Replaces BuildHamiltonianXXZForAbelianGroup (Q/exact_diagonalization/
// utils.cu:83-184). Row semantics are the reference's, entry by entry:
//   slot 0 = diagonal (column = the row itself), accumulated over terms whose
//   conjugated operators map the representative onto itself;
//   every other surviving term, in term order, gets its own slot (no merging):
//   column = row of the target's orbit, value = (coef * chi(gm)) * (n_t / n_s)
//   with gm the smallest element taking the orbit representative onto the
//   target, n = sqrt(G*m)/G the orbit norm (m = stabiliser size).
// Storage is SELL-64: rows in chunks of one wavefront, each chunk stored
// slot-major and padded to its own longest row, so an SpMV wave reads every
// slot with one coalesced load. Two passes: count -> chunk offsets -> fill.
// A packed build's fill writes 16-bit value codes instead of FP64 values; the
// codes in use are collected and turned into value indexes afterwards.
#include <cmath>
#include <cstdlib>
#include <memory>

#include "device_common.hpp"
#include "terms.hpp"

namespace qmb {

// 8 waves per block share the staged tables (rank, group, characters), so more
// waves fit a CU's LDS than with 4-wave blocks (6x6 ELL build: 145 ms with
// 4-wave blocks and the binomial rank loop, 131 ms with 8-wave blocks and the
// nibble rank table)
constexpr int kEllThreads = 512;

constexpr int kSetSlots = 4096;  // global set of distinct values
constexpr int kLdsSet = 256;     // per-block set
constexpr unsigned long long kEmptyKey = ~0ull;

__device__ inline uint32_t key_hash(unsigned long long k) {
    const uint32_t h = (uint32_t)(k >> 32) * 2654435761u;  // the value bits that vary most are the high ones
    return h ^ (h >> 15) ^ (uint32_t)k;
}

// Inserts key into an open-addressing set; false when the set is full.
__device__ inline bool set_insert(unsigned long long *set, int slots, unsigned long long key) {
    uint32_t h = key_hash(key) & (uint32_t)(slots - 1);
    for (int probe = 0; probe < slots; ++probe) {
        const unsigned long long cur = *(volatile unsigned long long *)(set + h);
        if (cur == key) return true;
        if (cur == kEmptyKey) {
            const unsigned long long old = atomicCAS(set + h, kEmptyKey, key);
            if (old == kEmptyKey || old == key) return true;
        }
        h = (h + 1) & (uint32_t)(slots - 1);
    }
    return false;
}

// Coded values (the build of a packed ELL). An off-diagonal value is
// (coef * chi(gm)) * (n(m_t) / n(m_s)): a term coefficient with the sign of its
// Sz factors, a character and two stabiliser sizes, each from a short list. The
// fill writes, instead of the FP64 value, the 16-bit code
//   (((coef class * 2 + sign) * n_chi + chi class) * n_m + m_s class) * n_m + m_t class
// and the diagonal straight into the FP64 diagonal array; the distinct values
// are formed afterwards from the codes in use, with the same arithmetic, so
// they are bit-identical to the FP64 form's.
constexpr int kMaxCodes = 1 << 14;

struct CodeView {
    const uint16_t *xcls = nullptr;  // [G] character class of each element
    const uint8_t *mcls = nullptr;   // [kMaxGroup + 1] class of each stabiliser size (0xFF: none)
    int n_chi = 0, n_m = 0;
    uint16_t *code = nullptr;        // [stored]
    double *diag = nullptr;          // [n_rows] (x 2 complex)
    int *bad = nullptr;              // set when a stabiliser size has no class
};

// Rank tables in dynamic LDS: the nibble table (rank_tab, n_nib > 0) when it
// fits, else the binomial table (binom). CODED (fill only): codes instead of
// values (CodeView).
template <class S, bool FILL, bool CPLX, bool CODED = false>
__global__ __launch_bounds__(kEllThreads) void ell_rows(GroupView gv, const uint64_t *binom,
                                                        const uint64_t *rank_tab, int n_nib, int n_c, int layout,
                                                        int64_t n_rows, const uint64_t *reps, const uint8_t *stab,
                                                        const uint8_t *nullf, const uint64_t *dict,
                                                        TermsView tv, const double2 *chi,
                                                        const double *nrm, int W, uint16_t *row_len, int *err,
                                                        int all_rows, const int64_t *chunk_ptr, int32_t *col,
                                                        double *val, CodeView cv) {
    extern __shared__ uint64_t rank_lds[];
    __shared__ uint8_t perm_lds[kMaxGroup * 64];
    __shared__ double2 s_chi[kMaxGroup];
    __shared__ double s_nrm[kMaxGroup + 1];
    __shared__ GroupLds gt;
    __shared__ uint16_t s_xcls[CODED ? kMaxGroup : 1];
    __shared__ uint8_t s_mcls[CODED ? kMaxGroup + 1 : 1];
    // the count pass over a sector without null orbits needs no orbit lookups
    // (row length = 1 + the acting off-diagonal terms): no tables to stage
    const bool lookups = FILL || !all_rows;
    if (lookups) {
        stage_perm(perm_lds, gv);
        stage_group(gt, gv);
        if (n_nib) {
            for (int i = threadIdx.x; i < n_nib * n_c * 16; i += blockDim.x) rank_lds[i] = rank_tab[i];
        } else {
            stage_binom(rank_lds, binom);
        }
        for (int i = threadIdx.x; i < gv.G; i += blockDim.x) s_chi[i] = chi[i];
        for (int i = threadIdx.x; i <= gv.G; i += blockDim.x) s_nrm[i] = nrm[i];
        if (CODED) {
            for (int i = threadIdx.x; i < gv.G; i += blockDim.x) s_xcls[i] = cv.xcls[i];
            for (int i = threadIdx.x; i <= kMaxGroup; i += blockDim.x) s_mcls[i] = cv.mcls[i];
        }
        __syncthreads();
    }
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const int64_t c = r / kChunk;
    const int lane = (int)(r % kChunk);
    int64_t base = 0;
    int L = 0;
    if (FILL) {
        base = chunk_ptr[c];
        L = (int)((chunk_ptr[c + 1] - base) / kChunk);
    }
    int slot = 1;
    const bool null_row = (layout == QMB_LAYOUT_REFERENCE) && nullf[r];
    double dr = 0.0, di = 0.0;
    if (!null_row) {
        const S src = (S)reps[r];
        const int m_src = FILL ? (int)stab[r] : 0;
        const double n_src = FILL ? s_nrm[m_src] : 0.0;
        // Terms are taken 64 at a time. Pass A applies the operators (cheap bit
        // tests, wave-uniform term index), adds diagonal terms in term order and
        // marks the lane's off-diagonal candidates; pass B pops each lane's
        // candidates in term order, so every iteration of the expensive orbit
        // lookup keeps almost all lanes busy.
        for (int h0 = 0; h0 < tv.n_terms; h0 += 64) {
            const int h1 = min(h0 + 64, tv.n_terms);
            uint64_t cand = 0;
            for (int h = h0; h < h1; ++h) {
                double cr, ci;
                S tgt;
                if (!apply_term<S>(tv, h, src, tgt, cr, ci)) continue;
                if (tgt == src) {
                    dr += cr;
                    di += ci;
                    continue;
                }
                cand |= 1ULL << (h - h0);
            }
            if (!FILL && all_rows) {
                // every orbit is a row (no null orbits for this irrep): the row
                // length needs no orbit lookup
                slot += __popcll(cand);
                continue;
            }
            while (cand) {
                const int h = h0 + __builtin_ctzll(cand);
                cand &= cand - 1;
                double cr, ci;
                S tgt;
                apply_term<S>(tv, h, src, tgt, cr, ci);
                int gm, mult;
                const S rep = gv.lut == 1   ? orbit_min_lut<S>(tgt, gv, gt, gm, mult)
                              : gv.lut == 2 ? orbit_min_ring<S>(tgt, gv, gt, gm, mult)
                              : gv.lut == 3 ? orbit_min_coset<S>(tgt, gv, gt, perm_lds, gm, mult)
                                            : orbit_min<S>(tgt, gv, perm_lds, gm, mult);
                const int64_t r2 = row_of_rank(
                    n_nib ? colex_rank_nib<S>(rep, rank_lds, n_nib, n_c) : colex_rank<S>(rep, rank_lds), dict);
                if (r2 < 0) continue;
                if (layout == QMB_LAYOUT_REFERENCE && nullf[r2]) continue;
                if (FILL && CODED) {
                    if (slot < L) {
                        const TermC &tm = tv.t[h];
                        const int neg = popc_s((S)(~src & (S)tm.sz)) & 1;
                        const int cs = s_mcls[m_src], ct = s_mcls[mult];
                        const bool ok = cs != 0xFF && ct != 0xFF;
                        if (!ok) *cv.bad = 1;
                        const int code =
                            ok ? ((((int)tm.vclass * 2 + neg) * cv.n_chi + s_xcls[gm]) * cv.n_m + cs) * cv.n_m + ct
                               : 0;
                        const int64_t pos = base + (int64_t)slot * kChunk + lane;
                        col[pos] = (int32_t)r2;
                        cv.code[pos] = (uint16_t)code;
                    }
                } else if (FILL) {
                    if (slot < L) {
                        // mult = the target row's stabiliser size (no read of stab[r2]);
                        // equal sizes give the ratio 1 exactly
                        const double ratio = mult == m_src ? 1.0 : s_nrm[mult] / n_src;
                        const double2 x = s_chi[gm];
                        const double pr = cr * x.x - ci * x.y;
                        const double pi = cr * x.y + ci * x.x;
                        const int64_t pos = base + (int64_t)slot * kChunk + lane;
                        col[pos] = (int32_t)r2;
                        if (CPLX) {
                            val[2 * pos] = pr * ratio;
                            val[2 * pos + 1] = pi * ratio;
                        } else {
                            val[pos] = pr * ratio;
                        }
                    }
                }
                ++slot;
            }
        }
    }
    const int len = null_row ? 0 : slot;
    if (!FILL) {
        if (len > W) atomicMax(err, 1);
        row_len[r] = (uint16_t)(len < 65535 ? len : 65535);
        return;
    }
    if (CODED) {
        // the diagonal into its own array; slot 0 and the padding slots hold the row
        if (CPLX) {
            cv.diag[2 * r] = null_row ? 0.0 : dr;
            cv.diag[2 * r + 1] = null_row ? 0.0 : di;
        } else {
            cv.diag[r] = null_row ? 0.0 : dr;
        }
        if (L > 0) col[base + lane] = (int32_t)r;
        for (int s = (len > 1 ? len : 1); s < L; ++s) col[base + (int64_t)s * kChunk + lane] = (int32_t)r;
        return;
    }
    // diagonal slot (also written for empty rows, like the reference's ind[id*W] = id)
    if (L > 0) {
        const int64_t pos0 = base + lane;
        col[pos0] = (int32_t)r;
        if (CPLX) {
            val[2 * pos0] = null_row ? 0.0 : dr;
            val[2 * pos0 + 1] = null_row ? 0.0 : di;
        } else {
            val[pos0] = null_row ? 0.0 : dr;
        }
    }
    for (int s = (len > 1 ? len : 1); s < L; ++s) {
        const int64_t pos = base + (int64_t)s * kChunk + lane;
        col[pos] = (int32_t)r;
        if (CPLX) {
            val[2 * pos] = 0.0;
            val[2 * pos + 1] = 0.0;
        } else {
            val[pos] = 0.0;
        }
    }
}

// Padding entries of the rows past n_rows in the last chunk.
__global__ void ell_pad_tail(int64_t n_rows, int64_t n_chunks, const int64_t *chunk_ptr, int32_t *col, double *val,
                             int cplx) {
    const int lane = threadIdx.x;
    const int64_t c = n_chunks - 1;
    const int64_t r = c * kChunk + lane;
    if (r < n_rows) return;
    const int64_t base = chunk_ptr[c];
    const int L = (int)((chunk_ptr[c + 1] - base) / kChunk);
    for (int s = 0; s < L; ++s) {
        const int64_t pos = base + (int64_t)s * kChunk + lane;
        col[pos] = 0;
        if (!val) continue;  // coded build: no FP64 values
        if (cplx) {
            val[2 * pos] = 0.0;
            val[2 * pos + 1] = 0.0;
        } else {
            val[pos] = 0.0;
        }
    }
}

// ---- packed values -----------------------------------------------------------
// The sector matrices of these models hold a few dozen distinct values (term
// scalars times characters times ratios of orbit norms), so a real ELL is
// repacked after the fill: every off-diagonal slot becomes one 32-bit word
// col | (index << col_bits) into a table of the distinct values, the diagonal
// moves to an FP64 array. The values and the order of the products are
// unchanged, so SpMV results are bit-identical; the matrix stream shrinks from
// 12 to 4 bytes per entry.
// Distinct off-diagonal values (slots 1 .. len-1 of every row): a per-block set
// in LDS, merged into the global set once per block; vflag[0] = 1 when a set
// fills. A separate streaming pass: inside the fill kernel, whose time goes to
// orbit lookups, the set probes cost more than re-reading the values.
// Complex values: the real and the imaginary parts go to two sets (vset and
// vset + kSetSlots).
template <bool CPLX>
__global__ __launch_bounds__(kEllThreads) void ell_collect_values(int64_t n_rows, const int64_t *chunk_ptr,
                                                                  const uint16_t *row_len, const double *val,
                                                                  unsigned long long *vset, int *vflag) {
    constexpr int kParts = CPLX ? 2 : 1;
    __shared__ unsigned long long set[kParts][kLdsSet];
    __shared__ int full;
    for (int i = threadIdx.x; i < kLdsSet; i += blockDim.x)
        for (int q = 0; q < kParts; ++q) set[q][i] = kEmptyKey;
    if (threadIdx.x == 0) full = 0;
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r / kChunk;
        const int lane = (int)(r % kChunk);
        const int64_t base = chunk_ptr[c];
        const int len = row_len[r];
        unsigned long long last[kParts];
        for (int q = 0; q < kParts; ++q) last[q] = kEmptyKey;
        for (int k = 1; k < len; ++k) {
            const int64_t p = base + (int64_t)k * kChunk + lane;
            for (int q = 0; q < kParts; ++q) {
                const unsigned long long key = (unsigned long long)__double_as_longlong(val[CPLX ? 2 * p + q : p]);
                if (key == last[q]) continue;
                last[q] = key;
                if (!set_insert(set[q], kLdsSet, key)) full = 1;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLdsSet; i += blockDim.x)
        for (int q = 0; q < kParts; ++q)
            if (set[q][i] != kEmptyKey && !set_insert(vset + q * kSetSlots, kSetSlots, set[q][i])) vflag[0] = 1;
    if (threadIdx.x == 0 && full) vflag[0] = 1;
}

// One block: the distinct values of the set in ascending bit order, +0.0 first
// (the padding value), into keys[0 .. n); meta = {n, ok}, ok = 0 when the set
// overflowed or n exceeds cap. The non-zero keys are gathered into LDS, then
// each goes to 1 + (number of smaller keys): the keys are distinct, so the
// ranks are a permutation.
__global__ __launch_bounds__(1024) void ell_sort_values(const unsigned long long *vset, const int *vflag,
                                                        unsigned long long *keys, int *meta, int cap) {
    __shared__ unsigned long long a[kSetSlots];
    __shared__ int m;
    if (threadIdx.x == 0) m = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < kSetSlots; i += blockDim.x) {
        const unsigned long long k = vset[i];
        if (k != kEmptyKey && k != 0ull) a[atomicAdd(&m, 1)] = k;
    }
    __syncthreads();
    const int n = m + 1;
    const bool ok = vflag[0] == 0 && n <= cap;
    if (ok) {
        for (int i = threadIdx.x; i < m; i += blockDim.x) {
            const unsigned long long k = a[i];
            int rank = 0;
            for (int j = 0; j < m; ++j) rank += a[j] < k ? 1 : 0;
            keys[1 + rank] = k;
        }
    }
    if (threadIdx.x == 0) {
        keys[0] = 0ull;
        meta[0] = n;
        meta[1] = ok ? 1 : 0;
    }
}

// In place, when meta[1] (ok): col[p] |= index << cb for the off-diagonal slots
// (index by binary search in the ascending keys), slot 0 -> the row itself,
// diag[r] = the old slot 0 value. *missing = 1 if a value is not in the table.
__global__ __launch_bounds__(kEllThreads) void ell_pack_values(int64_t n_rows, int64_t n_lanes,
                                                               const int64_t *chunk_ptr, const uint16_t *row_len,
                                                               uint32_t *colw, const double *val,
                                                               const unsigned long long *keys, const int *meta,
                                                               int cb, double *diag, int *missing) {
    __shared__ unsigned long long s_keys[kMaxValues];
    if (!meta[1]) return;
    const int n_keys = meta[0];
    for (int i = threadIdx.x; i < n_keys; i += blockDim.x) s_keys[i] = keys[i];
    const int top = n_keys > 1 ? 1 << (31 - __clz(n_keys - 1)) : 0;  // largest power of two <= n_keys - 1
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_lanes;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r / kChunk;
        const int lane = (int)(r % kChunk);
        const int64_t base = chunk_ptr[c];
        const int L = (int)((chunk_ptr[c + 1] - base) / kChunk);
        const int len = r < n_rows ? (int)row_len[r] : 0;
        if (r < n_rows) diag[r] = (len > 0 && L > 0) ? val[base + lane] : 0.0;
        if (L > 0) colw[base + lane] = (uint32_t)(r < n_rows ? r : 0);
        // four slots at a time, their searches interleaved: the last index whose
        // key is <= the value (keys[0] = 0 is <= every bit pattern), by a
        // branch-free descent over powers of two
        for (int k = 1; k < L; k += 4) {
            unsigned long long key[4];
            uint32_t w[4];
            int pos[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t p = base + (int64_t)(k + u) * kChunk + lane;
                w[u] = k + u < L ? colw[p] : 0u;
                key[u] = k + u < len ? (unsigned long long)__double_as_longlong(val[p]) : 0ull;
                pos[u] = 0;
            }
            for (int step = top; step > 0; step >>= 1) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = pos[u] + step;
                    if (q < n_keys && s_keys[q] <= key[u]) pos[u] = q;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k + u >= L) continue;
                if (k + u < len && s_keys[pos[u]] != key[u]) *missing = 1;
                colw[base + (int64_t)(k + u) * kChunk + lane] = w[u] | ((uint32_t)pos[u] << cb);
            }
        }
    }
}

__device__ inline uint32_t key_index(const unsigned long long *keys, int n, unsigned long long key, int *missing) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (keys[lo] != key) *missing = 1;
    return (uint32_t)lo;
}

// Complex values, in place when both tables fit (meta = {overflow, n_re, ok_re,
// missing, n_im, ok_im}) and the two indexes fit beside the column:
// col[p] |= (re index << cb) | (im index << (cb + re bits)); diag = the old slot 0.
__global__ __launch_bounds__(kEllThreads) void ell_pack_values_cplx(int64_t n_rows, int64_t n_lanes,
                                                                    const int64_t *chunk_ptr,
                                                                    const uint16_t *row_len, uint32_t *colw,
                                                                    const double *val,
                                                                    const unsigned long long *keys, const int *meta,
                                                                    int cb, double *diag, int *missing) {
    __shared__ unsigned long long s_re[kMaxValues], s_im[kMaxValues];
    const int n_re = meta[1], n_im = meta[4];
    if (!meta[2] || !meta[5]) return;
    const int rb = index_bits(n_re), sh2 = cb + rb;
    if (sh2 + index_bits(n_im) > 32) return;
    for (int i = threadIdx.x; i < n_re; i += blockDim.x) s_re[i] = keys[i];
    for (int i = threadIdx.x; i < n_im; i += blockDim.x) s_im[i] = keys[kMaxValues + i];
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_lanes;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = r / kChunk;
        const int lane = (int)(r % kChunk);
        const int64_t base = chunk_ptr[c];
        const int L = (int)((chunk_ptr[c + 1] - base) / kChunk);
        const int len = r < n_rows ? (int)row_len[r] : 0;
        if (r < n_rows) {
            const bool has = len > 0 && L > 0;
            diag[2 * r] = has ? val[2 * (base + lane)] : 0.0;
            diag[2 * r + 1] = has ? val[2 * (base + lane) + 1] : 0.0;
        }
        if (L > 0) colw[base + lane] = (uint32_t)(r < n_rows ? r : 0);
        for (int k = 1; k < L; ++k) {
            const int64_t p = base + (int64_t)k * kChunk + lane;
            uint64_t ir = 0, ii = 0;
            if (k < len) {
                ir = key_index(s_re, n_re, (unsigned long long)__double_as_longlong(val[2 * p]), missing);
                ii = key_index(s_im, n_im, (unsigned long long)__double_as_longlong(val[2 * p + 1]), missing);
            }
            colw[p] = colw[p] | (uint32_t)((ir << cb) | (ii << sh2));
        }
    }
}

// ---- coded values -------------------------------------------------------------
// Stabiliser sizes present among the rows: bit m of mask[0..4] (m <= kMaxGroup).
__global__ __launch_bounds__(kEllThreads) void ell_stab_sizes(int64_t n_rows, const uint8_t *stab, unsigned *mask) {
    __shared__ unsigned s[5];
    if (threadIdx.x < 5) s[threadIdx.x] = 0u;
    __syncthreads();
    uint64_t lo = 0, mid = 0;
    unsigned hi = 0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int m = stab[r];
        if (m < 64)
            lo |= 1ull << m;
        else if (m < 128)
            mid |= 1ull << (m - 64);
        else
            hi = 1u;
    }
    const unsigned w[5] = {(unsigned)lo, (unsigned)(lo >> 32), (unsigned)mid, (unsigned)(mid >> 32), hi};
    for (int i = 0; i < 5; ++i)
        if (w[i]) atomicOr(&s[i], w[i]);
    __syncthreads();
    if (threadIdx.x < 5 && s[threadIdx.x]) atomicOr(mask + threadIdx.x, s[threadIdx.x]);
}

// Codes in use (slots 1 .. len-1 of every row): bit c of used, through a
// per-block bitmap in LDS.
__global__ __launch_bounds__(kEllThreads) void ell_collect_codes(int64_t n_rows, const int64_t *chunk_ptr,
                                                                 const uint16_t *row_len, const uint16_t *code,
                                                                 int n_codes, unsigned *used) {
    __shared__ unsigned bits[kMaxCodes / 32];
    for (int i = threadIdx.x; i < kMaxCodes / 32; i += blockDim.x) bits[i] = 0u;
    __syncthreads();
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = chunk_ptr[r / kChunk] + r % kChunk;
        const int len = row_len[r];
        int last = -1;
        // eight slots' codes loaded at once, then marked (6x6 sector: 2.0 ms, 5.4 ms
        // one slot at a time)
        for (int k = 1; k < len; k += 8) {
            int c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) c[u] = k + u < len ? (int)code[base + (int64_t)(k + u) * kChunk] : -1;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (c[u] < 0 || c[u] == last || c[u] >= n_codes) continue;
                last = c[u];
                const unsigned b = 1u << (c[u] & 31);
                if (!(bits[c[u] >> 5] & b)) atomicOr(&bits[c[u] >> 5], b);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMaxCodes / 32; i += blockDim.x)
        if (bits[i]) atomicOr(used + i, bits[i]);
}

// The factors a code is made of (host tables, one device blob).
struct CodeTables {
    const double2 *coef = nullptr;  // [n_tc] term coefficient classes
    const double2 *chi = nullptr;   // [n_chi] character classes
    const int *m = nullptr;         // [n_m] stabiliser size of each class
    const double *nrm = nullptr;    // [G + 1] orbit norms
    int n_chi = 0, n_m = 0;
};

// One block: the value of every code in use, formed as the fill forms it
// (same operands, same operations, no contraction), into cval and the sets of
// distinct values (real parts in vset, imaginary parts in vset + kSetSlots).
__global__ __launch_bounds__(1024) void ell_code_values(const unsigned *used, int n_codes, CodeTables ct, int cplx,
                                                        double2 *cval, unsigned long long *vset, int *vflag) {
    for (int c = threadIdx.x; c < n_codes; c += blockDim.x) {
        if (!((used[c >> 5] >> (c & 31)) & 1u)) continue;
        int t = c;
        const int mt = t % ct.n_m;
        t /= ct.n_m;
        const int ms = t % ct.n_m;
        t /= ct.n_m;
        const int xc = t % ct.n_chi;
        t /= ct.n_chi;
        const double2 co = ct.coef[t >> 1];
        const bool neg = (t & 1) != 0;
        const double cr = neg ? -co.x : co.x;
        const double ci = neg ? -co.y : co.y;
        const int m_src = ct.m[ms], mult = ct.m[mt];
        const double n_src = ct.nrm[m_src];
        const double ratio = mult == m_src ? 1.0 : ct.nrm[mult] / n_src;
        const double2 x = ct.chi[xc];
        const double pr = cr * x.x - ci * x.y;
        const double pi = cr * x.y + ci * x.x;
        const double vr = pr * ratio, vi = pi * ratio;
        cval[c] = make_double2(vr, cplx ? vi : 0.0);
        if (!set_insert(vset, kSetSlots, (unsigned long long)__double_as_longlong(vr))) vflag[0] = 1;
        if (cplx && !set_insert(vset + kSetSlots, kSetSlots, (unsigned long long)__double_as_longlong(vi)))
            vflag[0] = 1;
    }
}

// One block, when the tables are complete (f: the packing flags): map[c] = the
// value index of code c (complex: re index | im index << 16).
__global__ __launch_bounds__(1024) void ell_code_map(const unsigned *used, int n_codes, const double2 *cval,
                                                     const unsigned long long *keys, int *f, int cplx,
                                                     uint32_t *map) {
    if (!f[2] || (cplx && !f[5])) return;
    for (int c = threadIdx.x; c < n_codes; c += blockDim.x) {
        if (!((used[c >> 5] >> (c & 31)) & 1u)) continue;
        const double2 v = cval[c];
        uint32_t m = key_index(keys, f[1], (unsigned long long)__double_as_longlong(v.x), f + 3);
        if (cplx) m |= key_index(keys + kMaxValues, f[4], (unsigned long long)__double_as_longlong(v.y), f + 3) << 16;
        map[c] = m;
    }
}

// In place, when the tables are complete and the indexes fit: every
// off-diagonal slot k < len gets its value index(es) beside the column (slot 0
// and the padding slots keep index 0, the value +0.0). Four slots at a time,
// each slot's map read issued with its word load (6x6 sector: 12.2 ms; eight
// slots, or the map staged in LDS and read after the loads, 14.4-16.8 ms).
__global__ __launch_bounds__(kEllThreads) void ell_pack_codes(int64_t n_rows, const int64_t *chunk_ptr,
                                                              const uint16_t *row_len, uint32_t *colw,
                                                              const uint16_t *code, const uint32_t *map,
                                                              const int *f, int cplx, int cb) {
    if (!f[2] || (cplx && !f[5])) return;
    int sh2 = 0;
    if (cplx) {
        sh2 = cb + index_bits(f[1]);
        if (sh2 + index_bits(f[4]) > 32) return;
    }
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t base = chunk_ptr[r / kChunk] + r % kChunk;
        const int len = row_len[r];
        for (int k = 1; k < len; k += 4) {
            uint32_t m[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t p = base + (int64_t)(k + u) * kChunk;
                m[u] = k + u < len ? map[code[p]] : 0u;
                w[u] = k + u < len ? colw[p] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k + u >= len) continue;
                const uint64_t ix = cplx ? (((uint64_t)(m[u] & 0xFFFFu) << cb) | ((uint64_t)(m[u] >> 16) << sh2))
                                         : (uint64_t)m[u] << cb;
                colw[base + (int64_t)(k + u) * kChunk] = w[u] | (uint32_t)ix;
            }
        }
    }
}

// Packing state of one build: the distinct values collected after the fill,
// then sort + pack on the device; the host reads the flags once at the end.
// flags: {overflow, n, ok, missing} (real) or {overflow, n_re, ok_re, missing,
// n_im, ok_im} (complex).
struct ValuePacking {
    bool on = false, cplx = false;
    int cb = 0, cap = 0;
    DevBuf vset, flags;
    // coded build (plan_codes, before the fill): flags[6] = a stabiliser size without a class
    bool coded = false;
    int n_codes = 0;
    DevBuf codes, used, cval, map, ctab;
    CodeView cv;
    CodeTables ct;
    bool prepare(qmb_ctx *ctx, const qmb_ell &e) {
        const int64_t n = e.n_rows;
        if (n < 2) return false;
        cplx = e.value_type == QMB_COMPLEX;
        cb = 1;
        while (((int64_t)1 << cb) < n) ++cb;
        if (cb > 31) return false;
        cap = cplx ? kMaxValues : (int)std::min<int64_t>(kMaxValues, (int64_t)1 << (32 - cb));
        const int parts = cplx ? 2 : 1;
        vset.alloc((size_t)parts * kSetSlots * 8);
        flags.alloc(8 * sizeof(int));
        QMB_HIP(hipMemsetAsync(vset.p, 0xFF, vset.bytes, ctx->stream));
        QMB_HIP(hipMemsetAsync(flags.p, 0, flags.bytes, ctx->stream));
        on = true;
        return true;
    }
    // Whether the fill writes codes: the coefficient classes of the terms, the
    // character classes of the irrep row and the stabiliser sizes present (one
    // small kernel and a read-back) must give at most kMaxCodes codes. Sets up
    // the code tables, the code array and the diagonal.
    // QMB_VALUE_CODES=0 in the environment selects the FP64 fill and the value
    // collection pass instead (the route of tables with too many codes), so
    // tests can compare the two routes.
    bool plan_codes(qmb_ctx *ctx, qmb_ell &e, const qmb_basis &b, const DeviceTerms &dt, const double *ch, int G) {
        const char *env = std::getenv("QMB_VALUE_CODES");
        if (env && env[0] == '0') return false;
        if (!on || dt.classes.size() > DeviceTerms::kMaxClasses || G > kMaxGroup) return false;
        DevBuf dmask(5 * sizeof(unsigned));
        QMB_HIP(hipMemsetAsync(dmask.p, 0, dmask.bytes, ctx->stream));
        const unsigned blocks = (unsigned)std::min<int64_t>(1024, (e.n_rows + kEllThreads - 1) / kEllThreads);
        hipLaunchKernelGGL(ell_stab_sizes, dim3(blocks), dim3(kEllThreads), 0, ctx->stream, e.n_rows,
                           b.stab.as<uint8_t>(), dmask.as<unsigned>());
        QMB_HIP(hipGetLastError());
        unsigned *hm = ctx->host_stage<unsigned>(5 * sizeof(unsigned));
        QMB_HIP(hipMemcpyAsync(hm, dmask.p, 5 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<int> m_of;
        std::vector<uint8_t> mcls(kMaxGroup + 1, 0xFF);
        for (int m = 0; m <= kMaxGroup; ++m)
            if ((hm[m >> 5] >> (m & 31)) & 1u) {
                if (m < 1 || m > G) return false;  // not a stabiliser size of this group
                mcls[m] = (uint8_t)m_of.size();
                m_of.push_back(m);
            }
        std::vector<double2> xc;
        std::vector<uint16_t> xcls(G);
        for (int g = 0; g < G; ++g) {
            const double2 v = make_double2(ch[2 * g], ch[2 * g + 1]);
            size_t k = 0;
            while (k < xc.size() && std::memcmp(&xc[k], &v, sizeof(v)) != 0) ++k;
            if (k == xc.size()) xc.push_back(v);
            xcls[g] = (uint16_t)k;
        }
        std::vector<double2> tc = dt.classes;
        if (tc.empty()) tc.push_back(make_double2(0.0, 0.0));
        const int64_t nc = (int64_t)tc.size() * 2 * (int64_t)xc.size() * (int64_t)m_of.size() * (int64_t)m_of.size();
        if (m_of.empty() || nc > kMaxCodes) return false;
        n_codes = (int)nc;
        // one blob: xcls, mcls, coefficient classes, character classes, sizes
        auto up16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
        const size_t o_m = up16((size_t)G * 2), o_tc = up16(o_m + kMaxGroup + 1), o_xc = up16(o_tc + tc.size() * 16),
                     o_mo = up16(o_xc + xc.size() * 16), total = up16(o_mo + m_of.size() * sizeof(int));
        std::vector<unsigned char> h(total, 0);
        std::memcpy(h.data(), xcls.data(), (size_t)G * 2);
        std::memcpy(h.data() + o_m, mcls.data(), mcls.size());
        std::memcpy(h.data() + o_tc, tc.data(), tc.size() * 16);
        std::memcpy(h.data() + o_xc, xc.data(), xc.size() * 16);
        std::memcpy(h.data() + o_mo, m_of.data(), m_of.size() * sizeof(int));
        ctab.alloc(total);
        QMB_HIP(hipMemcpy(ctab.p, h.data(), total, hipMemcpyHostToDevice));
        const char *d = static_cast<const char *>(ctab.p);
        codes.alloc((size_t)std::max<int64_t>(e.stored, 1) * sizeof(uint16_t));
        used.alloc(kMaxCodes / 8);
        QMB_HIP(hipMemsetAsync(used.p, 0, used.bytes, ctx->stream));
        cval.alloc((size_t)n_codes * sizeof(double2));
        map.alloc((size_t)n_codes * sizeof(uint32_t));
        e.diag.alloc((size_t)e.n_rows * (cplx ? 2 : 1) * sizeof(double));
        cv.xcls = reinterpret_cast<const uint16_t *>(d);
        cv.mcls = reinterpret_cast<const uint8_t *>(d + o_m);
        cv.n_chi = (int)xc.size();
        cv.n_m = (int)m_of.size();
        cv.code = codes.as<uint16_t>();
        cv.diag = e.diag.as<double>();
        cv.bad = flags.as<int>() + 6;
        ct.coef = reinterpret_cast<const double2 *>(d + o_tc);
        ct.chi = reinterpret_cast<const double2 *>(d + o_xc);
        ct.m = reinterpret_cast<const int *>(d + o_mo);
        ct.nrm = dt.nrm;
        ct.n_chi = cv.n_chi;
        ct.n_m = cv.n_m;
        coded = true;
        return true;
    }
    // after a coded fill: codes in use -> their values -> sorted tables -> code
    // to index map -> packed words
    void launch_coded(qmb_ctx *ctx, qmb_ell &e) {
        const int parts = cplx ? 2 : 1;
        e.vtab.alloc((size_t)parts * kMaxValues * 8);
        int *f = flags.as<int>();
        unsigned long long *keys = e.vtab.as<unsigned long long>();
        const unsigned blocks = (unsigned)std::min<int64_t>(4096, (e.n_rows + kEllThreads - 1) / kEllThreads);
        hipLaunchKernelGGL(ell_collect_codes, dim3(blocks), dim3(kEllThreads), 0, ctx->stream, e.n_rows,
                           e.chunk_ptr.as<int64_t>(), e.row_len.as<uint16_t>(), codes.as<uint16_t>(), n_codes,
                           used.as<unsigned>());
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(ell_code_values, dim3(1), dim3(1024), 0, ctx->stream, used.as<unsigned>(), n_codes, ct,
                           cplx ? 1 : 0, cval.as<double2>(), vset.as<unsigned long long>(), f);
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(ell_sort_values, dim3(1), dim3(1024), 0, ctx->stream, vset.as<unsigned long long>(), f,
                           keys, f + 1, cap);
        QMB_HIP(hipGetLastError());
        if (cplx) {
            hipLaunchKernelGGL(ell_sort_values, dim3(1), dim3(1024), 0, ctx->stream,
                               vset.as<unsigned long long>() + kSetSlots, f, keys + kMaxValues, f + 4, cap);
            QMB_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(ell_code_map, dim3(1), dim3(1024), 0, ctx->stream, used.as<unsigned>(), n_codes,
                           cval.as<double2>(), keys, f, cplx ? 1 : 0, map.as<uint32_t>());
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(ell_pack_codes, dim3(blocks), dim3(kEllThreads), 0, ctx->stream, e.n_rows,
                           e.chunk_ptr.as<int64_t>(), e.row_len.as<uint16_t>(), e.col.as<uint32_t>(),
                           codes.as<uint16_t>(), map.as<uint32_t>(), f, cplx ? 1 : 0, cb);
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipMemcpyAsync(ctx->host_stage<int>(8 * sizeof(int)), flags.p, 8 * sizeof(int),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    void launch(qmb_ctx *ctx, qmb_ell &e) {
        const int parts = cplx ? 2 : 1;
        e.vtab.alloc((size_t)parts * kMaxValues * 8);
        e.diag.alloc((size_t)e.n_rows * parts * sizeof(double));
        int *f = flags.as<int>();
        unsigned long long *keys = e.vtab.as<unsigned long long>();
        const unsigned cblocks = (unsigned)std::min<int64_t>(4096, (e.n_rows + kEllThreads - 1) / kEllThreads);
        hipLaunchKernelGGL(cplx ? ell_collect_values<true> : ell_collect_values<false>, dim3(cblocks),
                           dim3(kEllThreads), 0, ctx->stream, e.n_rows, e.chunk_ptr.as<int64_t>(),
                           e.row_len.as<uint16_t>(), e.val.as<double>(), vset.as<unsigned long long>(), f);
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(ell_sort_values, dim3(1), dim3(1024), 0, ctx->stream, vset.as<unsigned long long>(), f,
                           keys, f + 1, cap);
        QMB_HIP(hipGetLastError());
        if (cplx) {
            hipLaunchKernelGGL(ell_sort_values, dim3(1), dim3(1024), 0, ctx->stream,
                               vset.as<unsigned long long>() + kSetSlots, f, keys + kMaxValues, f + 4, cap);
            QMB_HIP(hipGetLastError());
        }
        const int64_t lanes = e.n_chunks * kChunk;
        const unsigned blocks = (unsigned)std::min<int64_t>(4096, (lanes + kEllThreads - 1) / kEllThreads);
        if (cplx)
            hipLaunchKernelGGL(ell_pack_values_cplx, dim3(blocks), dim3(kEllThreads), 0, ctx->stream, e.n_rows,
                               lanes, e.chunk_ptr.as<int64_t>(), e.row_len.as<uint16_t>(), e.col.as<uint32_t>(),
                               e.val.as<double>(), keys, f, cb, e.diag.as<double>(), f + 3);
        else
            hipLaunchKernelGGL(ell_pack_values, dim3(blocks), dim3(kEllThreads), 0, ctx->stream, e.n_rows, lanes,
                               e.chunk_ptr.as<int64_t>(), e.row_len.as<uint16_t>(), e.col.as<uint32_t>(),
                               e.val.as<double>(), keys, f + 1, cb, e.diag.as<double>(), f + 3);
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipMemcpyAsync(ctx->host_stage<int>(8 * sizeof(int)), flags.p, 8 * sizeof(int),
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    // after the stream has drained (the flags were copied into the context's
    // staging buffer at the end of launch)
    // Returns whether the matrix is packed; a coded build that is not must be
    // filled again with FP64 values.
    bool finish(qmb_ctx *ctx, qmb_ell &e) {
        const int *h = ctx->host_stage<int>(8 * sizeof(int));
        bool ok = h[2] != 0;
        if (cplx) ok = ok && h[5] != 0 && cb + index_bits(h[1]) + index_bits(h[4]) <= 32;
        QMB_REQUIRE(!ok || h[3] == 0, "qmb_ell_build: internal error while packing the values");
        QMB_REQUIRE(!coded || h[6] == 0, "qmb_ell_build: internal error (a stabiliser size without a value code)");
        codes.release();
        if (ok) {
            e.val.release();
            e.packed = 1;
            e.col_bits = cb;
            e.n_values = h[1];
            if (cplx) {
                e.re_bits = index_bits(h[1]);
                e.n_values_im = h[4];
            }
        } else {
            e.diag.release();
            e.vtab.release();
        }
        return ok;
    }
};

}  // namespace qmb

using namespace qmb;

double qmb_ell::spmv_bytes() const {
    const double sv = value_type == QMB_COMPLEX ? 16.0 : 8.0;
    return (double)nnz * (4.0 + sv) + (double)n_rows * (2.0 * sv) + (double)n_rows * 2.0;
}

int64_t qmb_ell::device_bytes() const {
    return (int64_t)(chunk_ptr.bytes + row_len.bytes + col.bytes + val.bytes + diag.bytes + vtab.bytes);
}

double qmb_ell::format_bytes() const {
    if (!packed) return spmv_bytes();
    // one word per off-diagonal entry, the diagonal, x read once, y written, row lengths
    const int64_t n_diag = basis && basis->layout == QMB_LAYOUT_REFERENCE ? n_rows - basis->n_null : n_rows;
    const double sv = value_type == QMB_COMPLEX ? 16.0 : 8.0;
    return (double)(nnz - n_diag) * 4.0 + (double)n_rows * (3.0 * sv + 2.0) +
           (double)(n_values + n_values_im) * 8.0;
}

namespace qmb {
// The rank lookup table of a basis for the ELL kernels: the nibble table on the
// device when it fits the LDS budget, else none (the kernels stage the binomial
// table). lds_bytes = the dynamic LDS of a launch that makes lookups.
struct RankTables {
    static constexpr size_t kMaxBytes = 48 * 1024;
    DevBuf buf;
    int n_nib = 0, n_c = 0;
    size_t lds_bytes = 64 * kBinJ * sizeof(uint64_t);
    explicit RankTables(const qmb_basis &b) {
        int nn = 0, nc = 0;
        const std::vector<uint64_t> t = rank_nibble_table(b.n_sites, b.n_up, nn, nc);
        const size_t bytes = t.size() * sizeof(uint64_t);
        if (bytes > kMaxBytes) return;
        buf.alloc(bytes);
        QMB_HIP(hipMemcpy(buf.p, t.data(), bytes, hipMemcpyHostToDevice));
        n_nib = nn;
        n_c = nc;
        lds_bytes = bytes;
    }
    const uint64_t *tab() const { return n_nib ? buf.as<uint64_t>() : nullptr; }
};
}  // namespace qmb

extern "C" {

qmb_status qmb_ell_build(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *t, int value_type, int max_terms,
                         qmb_ell **out) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(out, "qmb_ell_build: out is NULL");
        *out = nullptr;
        QMB_REQUIRE(ctx && b && t, "qmb_ell_build: NULL argument");
        QMB_REQUIRE(b->ctx == ctx, "qmb_ell_build: basis belongs to another context");
        QMB_REQUIRE(t->n_sites == b->n_sites, "qmb_ell_build: term table and basis disagree on n_sites");
        const bool keep_fp64 = (value_type & QMB_KEEP_FP64_VALUES) != 0;
        const bool locality = (value_type & QMB_LOCALITY_SLOTS) != 0;
        value_type &= ~(QMB_KEEP_FP64_VALUES | QMB_LOCALITY_SLOTS);
        QMB_REQUIRE(value_type == QMB_REAL || value_type == QMB_COMPLEX, "qmb_ell_build: bad value type");
        const qmb_group *g = b->group;
        const double *ch = &g->chars[(size_t)b->irrep * g->G * 2];
        if (value_type == QMB_REAL) {
            for (int e = 0; e < g->G; ++e)
                QMB_REQUIRE(ch[2 * e + 1] == 0.0, "qmb_ell_build: REAL values need a real character row");
            for (int h = 0; h < t->n_terms; ++h)
                QMB_REQUIRE(t->scalar[2 * h + 1] == 0.0, "qmb_ell_build: REAL values need real term scalars");
        }
        int W = max_terms;
        if (W <= 0) W = t->n_terms * b->n_up / b->n_sites + 1;  // reference rule
        QMB_REQUIRE(W <= 65535, "qmb_ell_build: max_terms too large");
        QMB_HIP(hipSetDevice(ctx->device));
        auto e = std::make_unique<qmb_ell>();
        e->ctx = ctx;
        e->basis = b;
        e->value_type = value_type;
        e->max_terms = W;
        e->n_rows = b->n_rows;
        e->n_chunks = (b->n_rows + kChunk - 1) / kChunk;
        // term table and irrep data on device
        const std::vector<int> order = locality ? locality_order(*t) : std::vector<int>();
        const DeviceTerms dt(*t, ch, g->G, locality ? &order : nullptr);
        const TermsView tv = dt.view;
        DevBuf d_err(sizeof(int));
        QMB_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int), ctx->stream));
        const GroupView gv = b->view();
        const int64_t n = b->n_rows;
        e->row_len.alloc((size_t)std::max<int64_t>(n, 1) * sizeof(uint16_t));
        e->chunk_ptr.alloc((size_t)(e->n_chunks + 1) * sizeof(int64_t));
        const RankTables rank(*b);
        const unsigned blocks = (unsigned)std::max<int64_t>(1, (n + kEllThreads - 1) / kEllThreads);
        const bool cplx = value_type == QMB_COMPLEX;
        if (n > 0) {
            const bool narrow = b->n_sites <= 32;
            auto kcount = narrow ? (cplx ? ell_rows<uint32_t, false, true> : ell_rows<uint32_t, false, false>)
                                 : (cplx ? ell_rows<uint64_t, false, true> : ell_rows<uint64_t, false, false>);
            // the count pass of a sector without null orbits makes no orbit lookups
            hipLaunchKernelGGL(kcount, dim3(blocks), dim3(kEllThreads), b->all_rows ? 0 : rank.lds_bytes,
                               ctx->stream, gv, ctx->binom.as<uint64_t>(), rank.tab(), rank.n_nib, rank.n_c,
                               b->layout, n, b->reps.as<uint64_t>(),
                               b->stab.as<uint8_t>(), b->nullf.as<uint8_t>(), b->dict.as<uint64_t>(),
                               tv, dt.chi, dt.nrm, W,
                               e->row_len.as<uint16_t>(), d_err.as<int>(), b->all_rows ? 1 : 0, nullptr, nullptr,
                               nullptr, CodeView{});
            QMB_HIP(hipGetLastError());
        }
        // the offsets are scanned before the overflow flag is read (one synchronisation)
        int herr = 0, max_len = 0;
        e->stored = n > 0 ? scan_chunks_i64(ctx, e->row_len.as<uint16_t>(), n, e->chunk_ptr.as<int64_t>(),
                                            e->n_chunks, &e->nnz, &max_len, d_err.as<int>(), &herr)
                          : 0;
        QMB_REQUIRE(herr == 0, "qmb_ell_build: a row needs more than max_terms = " + std::to_string(W) + " slots");
        e->max_row_len = max_len;
        e->col.alloc((size_t)std::max<int64_t>(e->stored, 1) * sizeof(int32_t));
        ValuePacking vp;
        if (n > 0 && !keep_fp64) vp.prepare(ctx, *e);
        // a packed build writes value codes (2 B per entry) instead of FP64 values
        // when the factors of the values give few enough codes
        const bool coded = vp.on && vp.plan_codes(ctx, *e, *b, dt, ch, g->G);
        // the fill (coded, or with FP64 values) and the padding of the last chunk
        auto fill = [&](bool with_codes) {
            const bool narrow = b->n_sites <= 32;
            auto kfill = with_codes
                             ? (narrow ? (cplx ? ell_rows<uint32_t, true, true, true> : ell_rows<uint32_t, true, false, true>)
                                       : (cplx ? ell_rows<uint64_t, true, true, true> : ell_rows<uint64_t, true, false, true>))
                             : (narrow ? (cplx ? ell_rows<uint32_t, true, true> : ell_rows<uint32_t, true, false>)
                                       : (cplx ? ell_rows<uint64_t, true, true> : ell_rows<uint64_t, true, false>));
            double *val = with_codes ? nullptr : e->val.as<double>();
            hipLaunchKernelGGL(kfill, dim3(blocks), dim3(kEllThreads), rank.lds_bytes, ctx->stream, gv,
                               ctx->binom.as<uint64_t>(), rank.tab(), rank.n_nib, rank.n_c, b->layout, n,
                               b->reps.as<uint64_t>(),
                               b->stab.as<uint8_t>(), b->nullf.as<uint8_t>(), b->dict.as<uint64_t>(),
                               tv, dt.chi, dt.nrm, W,
                               e->row_len.as<uint16_t>(), d_err.as<int>(), 0, e->chunk_ptr.as<int64_t>(),
                               e->col.as<int32_t>(), val, with_codes ? vp.cv : CodeView{});
            QMB_HIP(hipGetLastError());
            if (n % kChunk) {
                hipLaunchKernelGGL(ell_pad_tail, dim3(1), dim3(kChunk), 0, ctx->stream, n, e->n_chunks,
                                   e->chunk_ptr.as<int64_t>(), e->col.as<int32_t>(), val, cplx ? 1 : 0);
                QMB_HIP(hipGetLastError());
            }
        };
        if (!coded) e->val.alloc((size_t)std::max<int64_t>(e->stored, 1) * (cplx ? 16 : 8));
        if (n > 0) {
            fill(coded);
            if (coded)
                vp.launch_coded(ctx, *e);
            else if (vp.on)
                vp.launch(ctx, *e);
        }
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        if (vp.on && !vp.finish(ctx, *e) && coded) {
            // more distinct values than the index bits hold: the FP64 form
            e->val.alloc((size_t)std::max<int64_t>(e->stored, 1) * (cplx ? 16 : 8));
            fill(false);
            QMB_HIP(hipStreamSynchronize(ctx->stream));
        }
        *out = e.release();
    });
}

qmb_status qmb_ell_destroy(qmb_ell *e) {
    if (e && e->ctx) (void)hipStreamSynchronize(e->ctx->stream);
    qmb::release(e);
    return QMB_OK;
}

qmb_status qmb_ell_get_info(const qmb_ell *e, qmb_ell_info *info) {
    return guarded([&] {
        QMB_REQUIRE(e && info, "qmb_ell_get_info: NULL argument");
        info->n_rows = e->n_rows;
        info->nnz = e->nnz;
        info->stored = e->stored;
        info->max_row_len = e->max_row_len;
        info->max_terms = e->max_terms;
        info->value_type = e->value_type;
        info->chunk_rows = kChunk;
        info->device_bytes = e->device_bytes();
        info->spmv_bytes = e->spmv_bytes();
        info->packed = e->packed;
        info->n_values = e->n_values + e->n_values_im;
        info->format_bytes = e->format_bytes();
    });
}

qmb_status qmb_ell_export_rows(const qmb_ell *e, int64_t row0, int64_t n_rows, uint32_t *ind, double *val,
                               int16_t *track) {
    return guarded([&] {
        QMB_REQUIRE(e && (n_rows == 0 || (ind && val && track)), "qmb_ell_export_rows: NULL argument");
        QMB_REQUIRE(row0 >= 0 && n_rows >= 0 && row0 + n_rows <= e->n_rows, "qmb_ell_export_rows: bad row range");
        QMB_HIP(hipSetDevice(e->ctx->device));
        QMB_HIP(hipStreamSynchronize(e->ctx->stream));
        const int W = e->max_terms;
        const int vs = e->value_type == QMB_COMPLEX ? 2 : 1;
        std::fill(ind, ind + (size_t)n_rows * W, 0u);
        std::fill(val, val + (size_t)n_rows * W * vs, 0.0);
        std::fill(track, track + n_rows, (int16_t)0);
        if (n_rows == 0) return;
        // the rows' chunks are one contiguous slice of the SELL arrays
        const int64_t c0 = row0 / kChunk, c1 = (row0 + n_rows + kChunk - 1) / kChunk;
        std::vector<int64_t> cp(c1 - c0 + 1);
        std::vector<uint16_t> rl(n_rows);
        QMB_HIP(hipMemcpy(cp.data(), e->chunk_ptr.as<int64_t>() + c0, cp.size() * sizeof(int64_t),
                          hipMemcpyDeviceToHost));
        QMB_HIP(hipMemcpy(rl.data(), e->row_len.as<uint16_t>() + row0, n_rows * sizeof(uint16_t),
                          hipMemcpyDeviceToHost));
        const int64_t base = cp[0], len_all = cp.back() - cp[0];
        std::vector<int32_t> cl(len_all);
        std::vector<double> vl((size_t)len_all * vs);
        QMB_HIP(hipMemcpy(cl.data(), e->col.as<int32_t>() + base, cl.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (e->packed) {
            // words -> columns and values; slot 0 = the diagonal
            std::vector<double> vt((size_t)kMaxValues * vs), dg((size_t)n_rows * vs);
            QMB_HIP(hipMemcpy(vt.data(), e->vtab.p, vt.size() * sizeof(double), hipMemcpyDeviceToHost));
            QMB_HIP(hipMemcpy(dg.data(), e->diag.as<double>() + (size_t)row0 * vs, dg.size() * sizeof(double),
                              hipMemcpyDeviceToHost));
            const uint32_t mask = (1u << e->col_bits) - 1u, rmask = (1u << e->re_bits) - 1u;
            const int sh2 = e->col_bits + e->re_bits;
            for (int64_t i = 0; i < n_rows; ++i) {
                const int64_t r = row0 + i;
                const int64_t c = r / kChunk - c0, lane = r % kChunk;
                const int64_t q0 = cp[c] - base + lane;
                const int L = (int)((cp[c + 1] - cp[c]) / kChunk);
                if (L > 0) {
                    cl[q0] = (int32_t)r;
                    for (int q = 0; q < vs; ++q) vl[(size_t)q0 * vs + q] = dg[(size_t)i * vs + q];
                }
                // only the row's own slots: padding words past its length carry no value index
                const int len = std::min<int>(rl[i], L);
                for (int s = 1; s < len; ++s) {
                    const int64_t q = q0 + (int64_t)s * kChunk;
                    const uint32_t w = (uint32_t)cl[q];
                    cl[q] = (int32_t)(w & mask);
                    if (vs == 1) {
                        vl[q] = vt[w >> e->col_bits];
                    } else {
                        vl[2 * q] = vt[(w >> e->col_bits) & rmask];
                        vl[2 * q + 1] = vt[kMaxValues + (uint32_t)((uint64_t)w >> sh2)];
                    }
                }
            }
        } else {
            QMB_HIP(hipMemcpy(vl.data(), e->val.as<double>() + (size_t)base * vs, vl.size() * sizeof(double),
                              hipMemcpyDeviceToHost));
        }
        for (int64_t i = 0; i < n_rows; ++i) {
            const int64_t r = row0 + i;
            const int64_t c = r / kChunk - c0, lane = r % kChunk;
            ind[(size_t)i * W] = (uint32_t)r;
            const int len = std::min<int>(rl[i], (int)((cp[c + 1] - cp[c]) / kChunk));
            track[i] = (int16_t)len;
            for (int s = 0; s < len; ++s) {
                const int64_t pos = cp[c] - base + (int64_t)s * kChunk + lane;
                ind[(size_t)i * W + s] = (uint32_t)cl[pos];
                for (int q = 0; q < vs; ++q) val[((size_t)i * W + s) * vs + q] = vl[(size_t)pos * vs + q];
            }
        }
    });
}

qmb_status qmb_ell_export_reference(const qmb_ell *e, uint32_t *ind, double *val, int16_t *track) {
    if (!e) {
        qmb::set_error("qmb_ell_export_reference: NULL argument");
        return QMB_ERR_INVALID;
    }
    return qmb_ell_export_rows(e, 0, e->n_rows, ind, val, track);
}

}  // extern "C"
