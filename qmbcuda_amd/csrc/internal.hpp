// Internal declarations of libqmb (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
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

// ---- device buffer ----------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 8;
        QMB_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

constexpr int kMaxGroup = 128;
constexpr int kChunk = 64;  // rows per SELL chunk = one wavefront

// Device view of a symmetry group (all arrays in device memory).
struct GroupView {
    int n_sites;
    int G;
    int torus;       // 1: cyclic-translation fast path, 0: generic permutation
    int block;       // torus block width B
    uint64_t full;   // mask of n_sites bits
    const uint64_t *keep;  // [G] torus in-block no-wrap mask
    const uint8_t *small;  // [G] in-block rotation
    const uint8_t *big;    // [G] whole-word rotation
    const uint8_t *flip;   // [G]
    const uint8_t *perm;   // [G*n_sites] generic destination table
    const uint8_t *inv;    // [G]
    int lut;               // 1: the group is every (block shift, in-block shift[, flip]) of a torus;
                           // 2: every rotation[, flip] of a ring
    int lut_flip;          //    with a spin-flip half
    const uint8_t *lut_tab;  // [kLutBytes] rotation minima and element index table (see device_common.hpp)
};

constexpr int kLutBytes = 256 + 256 + 2 * 8 * 8;

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
    int device = 0;
    hipStream_t stream = nullptr;
    qmb::DevBuf binom;  // uint64 [64*33] C(p, j)
    ~qmb_ctx() {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        binom.release();
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
    int lut = 0, lut_flip = 0;     // 1: full torus product with block width <= 8; 2: full ring
    std::vector<uint8_t> lut_tab;  // [kLutBytes]
};

struct qmb_terms {
    int n_sites = 0, n_terms = 0, max_ops = 0;
    std::vector<int32_t> op_type, site, n_ops;
    std::vector<double> scalar;  // re, im pairs
};

struct qmb_basis : qmb::Counted {
    qmb::Ref<qmb_ctx> ctx;
    qmb::Ref<const qmb_group> group;
    int n_sites = 0, n_up = 0, irrep = 0, layout = 0;
    int64_t sector_dim = 0, n_words = 0, n_rows = 0, n_null = 0;
    qmb::DevBuf bitmap;  // uint64 [n_words]: bit = representative kept as a row
    qmb::DevBuf prefix;  // uint32 [n_words]: rows before the word
    qmb::DevBuf reps;    // uint64 [n_rows]
    qmb::DevBuf stab;    // uint8 [n_rows]: stabiliser size
    qmb::DevBuf nullf;   // uint8 [n_rows]: 1 if the orbit is null for the irrep
    // device copies of the group tables
    qmb::DevBuf g_keep, g_small, g_big, g_flip, g_perm, g_inv, g_lut;
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
    qmb::DevBuf col;        // int32 [stored]
    qmb::DevBuf val;        // double or double2 [stored]
    double spmv_bytes() const;
    int64_t device_bytes() const;
};

namespace qmb {

// Exclusive scans on device (3-phase). Values come from a functor.
int64_t scan_popcount_u32(qmb_ctx *ctx, const uint64_t *words, uint32_t *prefix, int64_t n);
int64_t scan_chunks_i64(qmb_ctx *ctx, const uint16_t *row_len, int64_t n_rows, int64_t *chunk_ptr,
                        int64_t n_chunks, int64_t *nnz_out, int *max_len_out);

void ensure_binom(qmb_ctx *ctx);

// Host tridiagonal solver (bisection + inverse iteration), lanczos_host.cpp.
// T: diag a[0..m-1], off-diagonal b[1..m-1] (b[i] couples i-1 and i).
// Computes the k lowest eigenvalues w[0..k-1] and, per eigenvalue, the last
// component of its unit eigenvector; z0 (optional, length m) gets the unit
// eigenvector of the lowest eigenvalue with z0[0] >= 0.
void tridiag_lowest(int m, const double *a, const double *b, int k, double *w, double *zlast,
                    double *z0);

}  // namespace qmb
