// Internal declarations of libqmb (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "qmb.h"

namespace qmb {

// ---- errors -----------------------------------------------------------------
void set_error(const std::string &msg);
qmb_status fail(qmb_status code, const std::string &msg);

struct Error {
    qmb_status code;
    std::string msg;
};

#define QMB_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::qmb::Error{e_ == hipErrorOutOfMemory ? QMB_ERR_OOM : QMB_ERR_HIP,    \
                               std::string(#call) + ": " + hipGetErrorString(e_)};       \
    } while (0)

#define QMB_REQUIRE(cond, msg)                                                           \
    do {                                                                                 \
        if (!(cond)) throw ::qmb::Error{QMB_ERR_INVALID, (msg)};                         \
    } while (0)

// Runs f, turning exceptions into a status + last-error text (no exception
// crosses the C ABI).
template <class F>
qmb_status guarded(F &&f) {
    try {
        f();
        return QMB_OK;
    } catch (const Error &e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(QMB_ERR_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(QMB_ERR_INVALID, e.what());
    }
}

// ---- device memory ------------------------------------------------------------
// A context caches the device blocks its buffers release and hands them out
// again: hipMalloc/hipFree cost 10-20 us each (about 40 per small solve), and
// memory freed to the driver is scrubbed before it can be mapped again, which
// made a repeated 6x6 solve (~170 GB of matrix, temporaries and kept Lanczos
// vectors) several times slower than the first. A released block is reused only
// once the context's stream has drained, so no queued kernel and no synchronous
// copy can touch a block that changed hands. Blocks of 1 GiB and more are reused
// for requests within 1/8 of their size, smaller ones within a factor 2; the
// cache holds at most max_cached bytes (3/4 of the device memory) and is
// emptied when an allocation fails, before it is retried.
struct DevPool {
    hipStream_t stream = nullptr;
    std::mutex m;
    std::multimap<size_t, void *> ready;    // free blocks by size
    std::vector<std::pair<size_t, void *>> pending;  // released while the stream may still use them
    size_t cached = 0;
    size_t max_cached = (size_t)16 << 30;  // set from the device memory when the context is created
    static constexpr size_t kBig = (size_t)1 << 30;

    void *get(size_t n, size_t &cap) {
        {
            std::lock_guard<std::mutex> lock(m);
            if (!pending.empty() && hipStreamQuery(stream) == hipSuccess) {
                for (auto &b : pending) ready.emplace(b.first, b.second);
                pending.clear();
            }
            auto it = ready.lower_bound(n);
            if (it != ready.end() && it->first <= (n >= kBig ? n + n / 8 : 2 * n)) {
                void *p = it->second;
                cap = it->first;
                cached -= cap;
                ready.erase(it);
                return p;
            }
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            trim();
            e = hipMalloc(&p, n);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{e == hipErrorOutOfMemory ? QMB_ERR_OOM : QMB_ERR_HIP,
                        std::string("hipMalloc: ") + hipGetErrorString(e)};
        }
        cap = n;
        return p;
    }
    size_t cached_bytes() {
        std::lock_guard<std::mutex> lock(m);
        return cached;
    }
    void put(void *p, size_t cap) {
        {
            std::lock_guard<std::mutex> lock(m);
            if (cached + cap <= max_cached) {
                pending.emplace_back(cap, p);
                cached += cap;
                return;
            }
        }
        (void)hipFree(p);
    }
    void trim() {
        std::lock_guard<std::mutex> lock(m);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &b : ready) (void)hipFree(b.second);
        for (auto &b : pending) (void)hipFree(b.second);
        ready.clear();
        pending.clear();
        cached = 0;
    }
    ~DevPool() { trim(); }
};

// The pool that buffers allocated on this thread draw from: set for the
// duration of a context's entry point (PoolScope); none -> plain hipMalloc.
DevPool *&current_pool();

struct PoolScope {
    DevPool *prev;
    explicit PoolScope(DevPool *p) : prev(current_pool()) { current_pool() = p; }
    ~PoolScope() { current_pool() = prev; }
    PoolScope(const PoolScope &) = delete;
    PoolScope &operator=(const PoolScope &) = delete;
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;  // requested size
    size_t cap = 0;    // block size
    DevPool *pool = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), cap(o.cap), pool(o.pool) {
        o.p = nullptr;
        o.bytes = o.cap = 0;
        o.pool = nullptr;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            cap = o.cap;
            pool = o.pool;
            o.p = nullptr;
            o.bytes = o.cap = 0;
            o.pool = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 8;
        pool = current_pool();
        if (pool) {
            p = pool->get(n, cap);
        } else {
            QMB_HIP(hipMalloc(&p, n));
            cap = n;
        }
        bytes = n;
    }
    void release() {
        if (p) {
            if (pool)
                pool->put(p, cap);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = cap = 0;
        pool = nullptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// Page-locked host memory (hipHostMalloc) for copies that must not block the host.
struct PinnedBuf {
    void *p = nullptr;
    explicit PinnedBuf(size_t n) { QMB_HIP(hipHostMalloc(&p, n == 0 ? 8 : n, hipHostMallocDefault)); }
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

constexpr int kMaxGroup = 128;
constexpr int kChunk = 64;  // rows per SELL chunk = one wavefront
constexpr int kMaxValues = 1024;  // value table of a packed ELL (per part for complex values)

// smallest b with 2^b >= n (n >= 1)
__host__ __device__ inline int index_bits(int n) {
    int b = 0;
    while ((1 << b) < n) ++b;
    return b;
}

// Device view of a symmetry group (all arrays in device memory).
struct GroupView {
    int n_sites;
    int G;
    int torus;       // 1: cyclic-translation fast path, 0: generic permutation
    int block;       // torus block width B
    uint64_t full;   // mask of n_sites bits
    uint64_t rep_unit;  // torus: bit 0 of every block (sum of 1 << (b * block))
    const uint64_t *keep;  // [G] torus in-block no-wrap mask
    const uint8_t *small;  // [G] in-block rotation
    const uint8_t *big;    // [G] whole-word rotation
    const uint8_t *flip;   // [G]
    const uint8_t *perm;   // [G*n_sites] generic destination table
    const uint8_t *inv;    // [G]
    int lut;               // 1: the group is every (block shift, in-block shift[, flip]) of a torus;
                           // 2: every rotation[, flip] of a ring
                           // 3: every rotation[, flip] of a ring times n_cos coset representatives
    int lut_flip;          //    with a spin-flip half
    const uint8_t *lut_tab;  // [kLutBytes] rotation minima and element index table (see device_common.hpp)
    int n_cos;
    int cos_rev;             // bit i: c_i is the site reversal s -> N-1-s
    const uint8_t *cos_tab;  // lut 3: [kCosBytes] element index of (rotation s, flip f) o c_i at
                             // i * 128 + f * 64 + s (c_i first; entry i * 128 is c_i itself)
};

constexpr int kLutBytes = 256 + 256 + 2 * 8 * 8;
constexpr int kMaxCosets = 8;
constexpr int kCosBytes = kMaxCosets * 128;

// Handles are reference counted: a child (basis, ELL, network) holds its
// parents (context, group, basis), so the C ABI's destroy calls may come in any
// order; the last holder frees the object.
struct Counted {
    mutable std::atomic<int> refs{1};
};

template <class T>
inline void retain(const T *p) {
    if (p) p->refs.fetch_add(1, std::memory_order_relaxed);
}

template <class T>
inline void release(const T *p) {
    if (p && p->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete p;
}

template <class T>
struct Ref {
    T *p = nullptr;
    Ref() = default;
    Ref(const Ref &) = delete;
    Ref &operator=(const Ref &) = delete;
    Ref(Ref &&o) noexcept : p(o.p) { o.p = nullptr; }
    Ref &operator=(Ref &&o) noexcept {
        if (this != &o) {
            release(p);
            p = o.p;
            o.p = nullptr;
        }
        return *this;
    }
    Ref &operator=(T *q) {
        retain(q);
        release(p);
        p = q;
        return *this;
    }
    ~Ref() { release(p); }
    T *operator->() const { return p; }
    operator T *() const { return p; }
};

}  // namespace qmb

struct qmb_ctx : qmb::Counted {
    qmb::DevPool pool;  // declared first: destroyed after every buffer of the context
    int device = 0;
    hipStream_t stream = nullptr;
    qmb::DevBuf binom;  // uint64 [64*33] C(p, j)
    // pinned host staging for small per-call read-backs (grows, kept)
    std::unique_ptr<qmb::PinnedBuf> stage;
    size_t stage_bytes = 0;
    template <class T>
    T *host_stage(size_t bytes) {
        if (bytes > stage_bytes) {
            stage.reset();
            stage_bytes = std::max<size_t>(bytes, 64 * 1024);
            stage = std::make_unique<qmb::PinnedBuf>(stage_bytes);
        }
        return stage->as<T>();
    }
    ~qmb_ctx() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        binom.release();
        stage.reset();
        pool.trim();
        pool.stream = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct qmb_group : qmb::Counted {
    int n_sites = 0, G = 0, n_irreps = 0;
    std::vector<uint32_t> perm;   // [G*n_sites]
    std::vector<uint8_t> flip;    // [G]
    std::vector<double> chars;    // [n_irreps*G*2]
    std::vector<int32_t> inv;     // [G]
    int torus = 0, block = 0;
    std::vector<uint64_t> keep;
    std::vector<uint8_t> small, big;
    std::vector<uint8_t> irrep_ok;  // [n_irreps]: the row is a one-dimensional representation
    uint64_t hash = 0;              // FNV-1a of the element tables, flips and characters
    int lut = 0, lut_flip = 0;     // 1: full torus product with block width <= 8; 2: full ring;
                                   // 3: a full ring subgroup H times coset representatives
    std::vector<uint8_t> lut_tab;  // [kLutBytes]
    std::vector<uint8_t> comp;     // [G*G]: index of a o b (b applied first)
    int n_cos = 0, cos_rev = 0;    // lut 3: cosets of the ring subgroup and their element table
    std::vector<uint8_t> cos_tab;  // [kCosBytes]
    std::vector<double> orbit_norms;  // [G + 1]: sqrt(G m)/G for stabiliser size m
};

struct qmb_terms {
    int n_sites = 0, n_terms = 0, max_ops = 0;
    std::vector<int32_t> op_type, site, n_ops;
    std::vector<double> scalar;  // re, im pairs
    std::vector<uint8_t> compiled;  // the terms compiled to masks (TermC[n_terms], host), made at creation
};

struct qmb_basis : qmb::Counted {
    qmb::Ref<qmb_ctx> ctx;
    qmb::Ref<const qmb_group> group;
    int n_sites = 0, n_up = 0, irrep = 0, layout = 0;
    int64_t sector_dim = 0, n_words = 0, n_rows = 0, n_null = 0;
    // rank -> row dictionary, one 16-byte record per 64 sector ranks:
    // dict[2w] = bit i set when rank 64w + i is a kept representative,
    // dict[2w + 1] = rows before the word (one random access per lookup)
    qmb::DevBuf dict;
    qmb::DevBuf reps;    // uint64 [n_rows]
    qmb::DevBuf stab;    // uint8 [n_rows]: stabiliser size
    qmb::DevBuf nullf;   // uint8 [n_rows]: 1 if the orbit is null for the irrep
    // device copies of the group tables
    // one device block: keep (8 B each), then small, big, flip, inv (1 B each),
    // perm (G * n_sites bytes), the lookup table and the coset table
    qmb::DevBuf g_tab;
    size_t o_small = 0, o_big = 0, o_flip = 0, o_inv = 0, o_perm = 0, o_lut = 0, o_cos = 0;
    qmb::GroupView view() const;
    uint64_t triv_lo = 0, triv_hi = 0;  // elements with character 1 for the irrep
    bool all_rows = false;              // every element has character 1: no null orbits
};

struct qmb_ell : qmb::Counted {
    qmb::Ref<qmb_ctx> ctx;
    qmb::Ref<const qmb_basis> basis;
    int value_type = QMB_REAL;
    int max_terms = 0;
    int64_t n_rows = 0, n_chunks = 0, nnz = 0, stored = 0;
    int max_row_len = 0;
    qmb::DevBuf chunk_ptr;  // int64 [n_chunks+1]
    qmb::DevBuf row_len;    // uint16 [n_rows]
    qmb::DevBuf col;        // int32 [stored]; packed: uint32 col | (value index << col_bits), slot 0 unused
    qmb::DevBuf val;        // double or double2 [stored]; empty when packed
    // Packed values: the matrix has few distinct values, so an off-diagonal slot
    // stores value indexes next to its column and the diagonal is kept in FP64
    // per row. Real: col | (index << col_bits) into vtab[0 .. n_values). Complex:
    // col | (re index << col_bits) | (im index << (col_bits + re_bits)) into
    // vtab[0 .. n_values) (real parts) and vtab[kMaxValues ..] (n_values_im
    // imaginary parts).
    int packed = 0;
    int col_bits = 0;
    int n_values = 0;
    int re_bits = 0, n_values_im = 0;  // complex only
    qmb::DevBuf diag;       // double [n_rows] or double2 [n_rows] (packed)
    qmb::DevBuf vtab;       // double [kMaxValues] (x 2 complex), ascending bit patterns, entry 0 = +0.0
    double spmv_bytes() const;    // algorithmic bytes of one SpMV (int32 column + FP64 value per entry)
    double format_bytes() const;  // bytes of one SpMV in the stored format
    int64_t device_bytes() const;
};

namespace qmb {

// A caller's buffer must be device memory of the context's GPU with `bytes`
// bytes inside one allocation, or QMB_ERR_INVALID (a host pointer handed to a
// kernel would fault the GPU).
inline void require_device_range(const qmb_ctx *ctx, const void *p, size_t bytes, const std::string &what) {
    hipPointerAttribute_t a{};
    hipError_t r = hipPointerGetAttributes(&a, p);
    if (r != hipSuccess) (void)hipGetLastError();
    QMB_REQUIRE(r == hipSuccess && (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) &&
                    a.device == ctx->device,
                what + " is not device memory of the context's GPU");
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    r = hipMemGetAddressRange(&base, &size, const_cast<void *>(p));
    if (r != hipSuccess) (void)hipGetLastError();
    QMB_REQUIRE(r == hipSuccess && static_cast<const char *>(p) + bytes <= static_cast<const char *>(base) + size,
                what + " is smaller than the data it must hold");
}

// Exclusive scans on device (3-phase). Values come from a functor.
int64_t scan_popcount_dict(qmb_ctx *ctx, uint64_t *dict, int64_t n_words);
// Chunk offsets of a SELL-64 layout from its row lengths (chunk_ptr[n_chunks] =
// stored total, returned), with the entry count and longest row; d_err (if set)
// is read back in the same synchronisation into err_out.
int64_t scan_chunks_i64(qmb_ctx *ctx, const uint16_t *row_len, int64_t n_rows, int64_t *chunk_ptr,
                        int64_t n_chunks, int64_t *nnz_out, int *max_len_out, const int *d_err = nullptr,
                        int *err_out = nullptr);

void ensure_binom(qmb_ctx *ctx);

// Host tridiagonal solver (bisection + inverse iteration), lanczos_host.cpp.
// T: diag a[0..m-1], off-diagonal b[1..m-1] (b[i] couples i-1 and i).
// Computes the k lowest eigenvalues w[0..k-1] and, per eigenvalue, the last
// component of its unit eigenvector; z0 (optional, length m) gets the unit
// eigenvector of the lowest eigenvalue with z0[0] >= 0.
void tridiag_lowest(int m, const double *a, const double *b, int k, double *w, double *zlast,
                    double *z0);
// The k highest eigenvalues of T, largest first, and the last component of each
// unit eigenvector (up to sign).
void tridiag_highest(int m, const double *a, const double *b, int k, double *w, double *zlast);
// Unit eigenvector z[0..m) of T for its eigenvalue lam (inverse iteration).
void tridiag_vector(int m, const double *a, const double *b, double lam, double *z);

}  // namespace qmb
