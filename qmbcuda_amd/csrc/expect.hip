// Expectation values of Sz-conserving operators in a symmetry-reduced sector.
//
// Replaces the device part of ComputeStaticExpValZeroT
// (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:421-601):
// the kernel ExpecValArbOperatorAbelianIrrepProjHeisenbergZeroT
// (Q/exact_diagonalization/utils.cu:451-548) and its thrust::reduce.
//
// Rows of the sector matrix represent the states v_r = sum_g conj(chi(g)) g|rep_r>
// (normalised); an eigenvector x of the sector matrix is the many-body state
// Psi = sum_r conj(x_r) v_r. For any operator A with a conjugated term table,
// <Psi|A|Psi> = conj(sum_r conj(x_r) (M x)_r), where M is the sector matrix of the
// group average (1/G) sum_g g A g^-1 (it has the same expectation value as A in
// every state of the sector). Row r of M x, for the element images u = j.rep_r:
//   sum_j sum_terms c(u -> u') conj(chi(j)) chi(gm(u')) (n_t / n_r) x_t / G,
// t = row of the orbit of u', gm(u') = the element taking its representative onto
// u' (the ELL kernel's formula, applied to every image of the source). When A
// commutes with the group the average is A itself and only j = identity is taken.
//
// Two deliberate differences from the reference kernel: it stops the loop over
// the images of a term at the first image the term annihilates (utils.cu:523,
// `if (interr_bool) break;`), dropping the remaining images of S+/S- terms, and it
// accumulates conj(phi_r) conj(V) phi_t, which equals <Psi|A|Psi> only when the
// vector is real up to a phase. Both agree with this kernel for Sz-diagonal
// operators and real vectors.
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "device_common.hpp"
#include "reduce.hpp"
#include "terms.hpp"

namespace qmb {

constexpr int kExpThreads = 256;
constexpr int kExpMaxBlocks = 4096;

template <class S, bool CPLX, bool AVG>
__global__ __launch_bounds__(kExpThreads) void expect_rows(GroupView gv, const uint64_t *binom, int layout,
                                                           int64_t n_rows, const uint64_t *reps, const uint8_t *stab,
                                                           const uint8_t *nullf, const uint64_t *dict,
                                                           TermsView tv, const double2 *chi,
                                                           const double *nrm, const double *x, double *partials) {
    __shared__ uint8_t perm_lds[kMaxGroup * 64];
    __shared__ uint64_t bin[64 * kBinJ];
    __shared__ double2 s_chi[kMaxGroup];
    __shared__ double s_nrm[kMaxGroup + 1];
    __shared__ GroupLds gt;
    __shared__ double red[kWaves];
    stage_perm(perm_lds, gv);
    stage_group(gt, gv);
    stage_binom(bin, binom);
    for (int i = threadIdx.x; i < gv.G; i += blockDim.x) s_chi[i] = chi[i];
    for (int i = threadIdx.x; i <= gv.G; i += blockDim.x) s_nrm[i] = nrm[i];
    __syncthreads();
    const double inv_g = AVG ? 1.0 / (double)gv.G : 1.0;
    double out_r = 0.0, out_i = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n_rows;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (layout == QMB_LAYOUT_REFERENCE && nullf[r]) continue;
        const S rep = (S)reps[r];
        const double n_src = s_nrm[stab[r]];
        double ar = 0.0, ai = 0.0;  // (M x)_r
        const int n_img = AVG ? gv.G : 1;
        for (int j = 0; j < n_img; ++j) {
            const S u = AVG ? apply_elem<S>(rep, j, gv, perm_lds) : rep;
            const double2 cj = AVG ? make_double2(s_chi[j].x, -s_chi[j].y) : make_double2(1.0, 0.0);
            for (int h0 = 0; h0 < tv.n_terms; h0 += 64) {
                const int h1 = min(h0 + 64, tv.n_terms);
                uint64_t cand = 0;
                double dr = 0.0, di = 0.0;
                for (int h = h0; h < h1; ++h) {
                    double cr, ci;
                    S tgt;
                    if (!apply_term<S>(tv, h, u, tgt, cr, ci)) continue;
                    if (tgt == u) {  // same orbit element: conj(chi(j)) chi(gm) = 1, n_t = n_r
                        dr += cr;
                        di += ci;
                    } else {
                        cand |= 1ULL << (h - h0);
                    }
                }
                if (CPLX) {
                    const double xr = x[2 * r], xi = x[2 * r + 1];
                    ar += dr * xr - di * xi;
                    ai += dr * xi + di * xr;
                } else {
                    const double xr = x[r];
                    ar += dr * xr;
                    ai += di * xr;
                }
                while (cand) {
                    const int h = h0 + __builtin_ctzll(cand);
                    cand &= cand - 1;
                    double cr, ci;
                    S tgt;
                    apply_term<S>(tv, h, u, tgt, cr, ci);
                    int gm, mult;
                    const S rep2 = gv.lut == 1   ? orbit_min_lut<S>(tgt, gv, gt, gm, mult)
                                   : gv.lut == 2 ? orbit_min_ring<S>(tgt, gv, gt, gm, mult)
                                   : gv.lut == 3 ? orbit_min_coset<S>(tgt, gv, gt, perm_lds, gm, mult)
                                                 : orbit_min<S>(tgt, gv, perm_lds, gm, mult);
                    const int64_t r2 = row_of_rep<S>(rep2, bin, dict);
                    if (r2 < 0) continue;
                    if (layout == QMB_LAYOUT_REFERENCE && nullf[r2]) continue;
                    const double ratio = s_nrm[mult] / n_src;  // mult = the target row's stabiliser size
                    const double2 cg = s_chi[gm];
                    // f = c * conj(chi(j)) * chi(gm) * ratio
                    const double qr = cj.x * cg.x - cj.y * cg.y, qi = cj.x * cg.y + cj.y * cg.x;
                    const double fr = (cr * qr - ci * qi) * ratio, fi = (cr * qi + ci * qr) * ratio;
                    if (CPLX) {
                        const double xr = x[2 * r2], xi = x[2 * r2 + 1];
                        ar += fr * xr - fi * xi;
                        ai += fr * xi + fi * xr;
                    } else {
                        const double xr = x[r2];
                        ar += fr * xr;
                        ai += fi * xr;
                    }
                }
            }
        }
        ar *= inv_g;
        ai *= inv_g;
        // x_r conj((M x)_r)
        const double xr = CPLX ? x[2 * r] : x[r], xi = CPLX ? x[2 * r + 1] : 0.0;
        out_r += xr * ar + xi * ai;
        out_i += xi * ar - xr * ai;
    }
    const double sr = block_sum(out_r, red);
    const double si = block_sum(out_i, red);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sr;
        partials[2 * blockIdx.x + 1] = si;
    }
}

// True when the operator commutes with every element: the images g A g^-1 (sites
// moved by the element, S+ <-> S- and Sz -> -Sz under a spin flip) have the same
// terms with the same summed scalars. Operators on one site keep their order.
// A term is keyed by its (site, op) byte pairs after a stable sort by site.
static bool commutes_with_group(const qmb_terms &t, const qmb_group &g) {
    using Entry = std::pair<std::string, std::complex<double>>;
    auto canon = [&](int elem) {
        std::vector<Entry> v;
        v.reserve(t.n_terms);
        const int n = t.n_sites;
        std::vector<std::pair<int, int>> ops;
        for (int h = 0; h < t.n_terms; ++h) {
            const int k = t.n_ops[h];
            ops.resize(k);
            std::complex<double> c(t.scalar[2 * h], t.scalar[2 * h + 1]);
            for (int q = 0; q < k; ++q) {
                int ty = t.op_type[(size_t)h * t.max_ops + q];
                const int s = t.site[(size_t)h * t.max_ops + q];
                const int d = elem < 0 ? s : (int)g.perm[(size_t)elem * n + s];
                if (elem >= 0 && g.flip[elem]) {
                    if (ty == QMB_OP_SP)
                        ty = QMB_OP_SM;
                    else if (ty == QMB_OP_SM)
                        ty = QMB_OP_SP;
                    else
                        c = -c;
                }
                ops[q] = {d, ty};
            }
            std::stable_sort(ops.begin(), ops.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
            std::string key(2 * k, '\0');
            for (int q = 0; q < k; ++q) {
                key[2 * q] = (char)ops[q].first;
                key[2 * q + 1] = (char)ops[q].second;
            }
            v.emplace_back(std::move(key), c);
        }
        std::sort(v.begin(), v.end(), [](const Entry &a, const Entry &b) { return a.first < b.first; });
        std::vector<Entry> merged;  // equal products summed
        for (auto &e : v) {
            if (!merged.empty() && merged.back().first == e.first)
                merged.back().second += e.second;
            else
                merged.push_back(std::move(e));
        }
        return merged;
    };
    const auto base = canon(-1);
    double scale = 0.0;
    for (const auto &e : base) scale = std::max(scale, std::abs(e.second));
    const double tol = 1e-12 * std::max(scale, 1e-300);
    auto same = [&](const std::vector<Entry> &a, const std::vector<Entry> &b) {
        size_t i = 0, j = 0;
        while (i < a.size() || j < b.size()) {
            if (j == b.size() || (i < a.size() && a[i].first < b[j].first)) {
                if (std::abs(a[i].second) > tol) return false;
                ++i;
            } else if (i == a.size() || b[j].first < a[i].first) {
                if (std::abs(b[j].second) > tol) return false;
                ++j;
            } else {
                if (std::abs(a[i].second - b[j].second) > tol) return false;
                ++i;
                ++j;
            }
        }
        return true;
    };
    for (int e = 0; e < g.G; ++e)
        if (!same(base, canon(e))) return false;
    return true;
}

// ---- the entry point that allocates nothing ----------------------------------
// Scratch of qmb_expectation_noalloc: one file-scope device array per GPU, held
// until the process ends. Layout: [0, 256) the result (re, im); [256, 256 +
// 64 KiB) the per-block partial sums; then the compiled terms, the character row
// (G double2) and the orbit-norm table (G + 1 doubles), each 256-byte aligned.
// Host code never uses the array's name as an address: the device address comes
// from hipGetSymbolAddress, once per device.
constexpr size_t kScratchBytes = (size_t)1 << 20;
constexpr size_t kScratchPart = 256;
constexpr size_t kScratchTerms = kScratchPart + (size_t)kExpMaxBlocks * 2 * sizeof(double);
constexpr size_t kScratchTabBytes = 4096;  // chi (<= 128 x 16 B) + norms (<= 129 x 8 B), aligned
constexpr int kNoAllocMaxTerms = (int)((kScratchBytes - kScratchTerms - kScratchTabBytes) / sizeof(TermC));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on `device` (the current device), looked up once per device.
static unsigned char *scratch_ptr(int device) {
    static std::mutex mu;
    static unsigned char *ptrs[kMaxDevices] = {};
    QMB_REQUIRE(device >= 0 && device < kMaxDevices, "qmb_expectation_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[device]) {
        void *p = nullptr;
        QMB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QMB_REQUIRE(p != nullptr, "qmb_expectation_noalloc: no device address for the scratch array");
        ptrs[device] = static_cast<unsigned char *>(p);
    }
    return ptrs[device];
}

// The partial sums in block order, as the host loop of qmb_expectation adds them.
__global__ __launch_bounds__(64) void expect_finish(const double *__restrict__ partials, int blocks,
                                                    double *__restrict__ out) {
    if (threadIdx.x != 0) return;
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < blocks; ++i) {
        sr += partials[2 * i];
        si += partials[2 * i + 1];
    }
    out[0] = sr;
    out[1] = si;
}

}  // namespace qmb

using namespace qmb;

extern "C" {

qmb_status qmb_expectation_noalloc(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *a, int is_complex,
                                   const double *x_dev, int flags, double *out, int *averaged) {
    return guarded([&] {
        QMB_REQUIRE(ctx && b && a && x_dev && out, "qmb_expectation_noalloc: NULL argument");
        QMB_REQUIRE(b->ctx == ctx, "qmb_expectation_noalloc: basis belongs to another context");
        QMB_REQUIRE(a->n_sites == b->n_sites, "qmb_expectation_noalloc: term table and basis disagree on n_sites");
        QMB_REQUIRE(is_complex == 0 || is_complex == 1, "qmb_expectation_noalloc: is_complex must be 0 or 1");
        QMB_REQUIRE(flags == 0 || flags == QMB_EXPECT_GROUP_AVERAGE, "qmb_expectation_noalloc: unknown flags");
        QMB_REQUIRE(a->n_terms <= kNoAllocMaxTerms,
                    "qmb_expectation_noalloc: " + std::to_string(a->n_terms) + " terms exceed the scratch limit of " +
                        std::to_string(kNoAllocMaxTerms) + " (use qmb_expectation)");
        const qmb_group *g = b->group;
        QMB_REQUIRE((size_t)g->G * 16 + 256 + (size_t)(g->G + 1) * 8 <= kScratchTabBytes,
                    "qmb_expectation_noalloc: group too large for the scratch tables");
        const double *ch = &g->chars[(size_t)b->irrep * g->G * 2];
        if (!is_complex)
            for (int e = 0; e < g->G; ++e)
                QMB_REQUIRE(ch[2 * e + 1] == 0.0, "qmb_expectation_noalloc: a real vector needs a real character row");
        const bool avg = (flags & QMB_EXPECT_GROUP_AVERAGE) || !commutes_with_group(*a, *g);
        if (averaged) *averaged = avg ? 1 : 0;
        out[0] = 0.0;
        out[1] = 0.0;
        const int64_t n = b->n_rows;
        if (n == 0 || a->n_terms == 0) return;
        QMB_HIP(hipSetDevice(ctx->device));
        QMB_REQUIRE(ctx->binom.p != nullptr, "qmb_expectation_noalloc: context has no binomial table");
        require_device_range(ctx, x_dev, (size_t)n * (is_complex ? 16 : 8), "qmb_expectation_noalloc: x_dev");
        unsigned char *sp = scratch_ptr(ctx->device);
        // tables into the scratch (host copies made when the handles were created)
        const int nt = a->n_terms;
        auto up256 = [](size_t v) { return (v + 255) & ~(size_t)255; };
        const size_t o_chi = kScratchTerms + up256((size_t)nt * sizeof(TermC));
        const size_t o_nrm = o_chi + up256((size_t)g->G * 16);
        QMB_HIP(hipMemcpyAsync(sp + kScratchTerms, a->compiled.data(), (size_t)nt * sizeof(TermC),
                               hipMemcpyHostToDevice, ctx->stream));
        QMB_HIP(hipMemcpyAsync(sp + o_chi, ch, (size_t)g->G * 16, hipMemcpyHostToDevice, ctx->stream));
        QMB_HIP(hipMemcpyAsync(sp + o_nrm, g->orbit_norms.data(), (size_t)(g->G + 1) * sizeof(double),
                               hipMemcpyHostToDevice, ctx->stream));
        const TermsView tv{nt, reinterpret_cast<const TermC *>(sp + kScratchTerms)};
        const double2 *chi = reinterpret_cast<const double2 *>(sp + o_chi);
        const double *nrm = reinterpret_cast<const double *>(sp + o_nrm);
        double *part = reinterpret_cast<double *>(sp + kScratchPart);
        double *res = reinterpret_cast<double *>(sp);
        const int blocks = (int)std::min<int64_t>(kExpMaxBlocks, (n + kExpThreads - 1) / kExpThreads);
        const bool narrow = b->n_sites <= 32;
        using K = void (*)(GroupView, const uint64_t *, int, int64_t, const uint64_t *, const uint8_t *,
                           const uint8_t *, const uint64_t *, TermsView, const double2 *,
                           const double *, const double *, double *);
        K k;
        if (narrow)
            k = is_complex ? (avg ? expect_rows<uint32_t, true, true> : expect_rows<uint32_t, true, false>)
                           : (avg ? expect_rows<uint32_t, false, true> : expect_rows<uint32_t, false, false>);
        else
            k = is_complex ? (avg ? expect_rows<uint64_t, true, true> : expect_rows<uint64_t, true, false>)
                           : (avg ? expect_rows<uint64_t, false, true> : expect_rows<uint64_t, false, false>);
        hipLaunchKernelGGL(k, dim3(blocks), dim3(kExpThreads), 0, ctx->stream, b->view(), ctx->binom.as<uint64_t>(),
                           b->layout, n, b->reps.as<uint64_t>(), b->stab.as<uint8_t>(), b->nullf.as<uint8_t>(),
                           b->dict.as<uint64_t>(), tv, chi, nrm, x_dev, part);
        QMB_HIP(hipGetLastError());
        hipLaunchKernelGGL(expect_finish, dim3(1), dim3(64), 0, ctx->stream, part, blocks, res);
        QMB_HIP(hipGetLastError());
        QMB_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
    });
}

qmb_status qmb_expectation(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *a, int is_complex, const double *x,
                           int x_on_device, int flags, double *out, int *averaged) {
    return guarded([&] {
        PoolScope pool_scope(ctx ? &ctx->pool : nullptr);
        QMB_REQUIRE(ctx && b && a && x && out, "qmb_expectation: NULL argument");
        QMB_REQUIRE(b->ctx == ctx, "qmb_expectation: basis belongs to another context");
        QMB_REQUIRE(a->n_sites == b->n_sites, "qmb_expectation: term table and basis disagree on n_sites");
        QMB_REQUIRE(is_complex == 0 || is_complex == 1, "qmb_expectation: is_complex must be 0 or 1");
        QMB_REQUIRE(flags == 0 || flags == QMB_EXPECT_GROUP_AVERAGE, "qmb_expectation: unknown flags");
        const qmb_group *g = b->group;
        const double *ch = &g->chars[(size_t)b->irrep * g->G * 2];
        if (!is_complex)
            for (int e = 0; e < g->G; ++e)
                QMB_REQUIRE(ch[2 * e + 1] == 0.0, "qmb_expectation: a real vector needs a real character row");
        const bool avg = (flags & QMB_EXPECT_GROUP_AVERAGE) || !commutes_with_group(*a, *g);
        if (averaged) *averaged = avg ? 1 : 0;
        out[0] = 0.0;
        out[1] = 0.0;
        const int64_t n = b->n_rows;
        if (n == 0 || a->n_terms == 0) return;
        QMB_HIP(hipSetDevice(ctx->device));
        const size_t xbytes = (size_t)n * (is_complex ? 16 : 8);
        DevBuf d_x;
        const double *xd = x;
        if (!x_on_device) {
            d_x.alloc(xbytes);
            QMB_HIP(hipMemcpyAsync(d_x.p, x, xbytes, hipMemcpyHostToDevice, ctx->stream));
            xd = d_x.as<double>();
        } else {
            require_device_range(ctx, x, xbytes, "qmb_expectation: x");
        }
        const DeviceTerms dt(*a, ch, g->G);
        const int blocks = (int)std::min<int64_t>(kExpMaxBlocks, (n + kExpThreads - 1) / kExpThreads);
        DevBuf d_part((size_t)blocks * 2 * sizeof(double));
        const bool narrow = b->n_sites <= 32;
        using K = void (*)(GroupView, const uint64_t *, int, int64_t, const uint64_t *, const uint8_t *,
                           const uint8_t *, const uint64_t *, TermsView, const double2 *,
                           const double *, const double *, double *);
        K k;
        if (narrow)
            k = is_complex ? (avg ? expect_rows<uint32_t, true, true> : expect_rows<uint32_t, true, false>)
                           : (avg ? expect_rows<uint32_t, false, true> : expect_rows<uint32_t, false, false>);
        else
            k = is_complex ? (avg ? expect_rows<uint64_t, true, true> : expect_rows<uint64_t, true, false>)
                           : (avg ? expect_rows<uint64_t, false, true> : expect_rows<uint64_t, false, false>);
        hipLaunchKernelGGL(k, dim3(blocks), dim3(kExpThreads), 0, ctx->stream, b->view(), ctx->binom.as<uint64_t>(),
                           b->layout, n, b->reps.as<uint64_t>(), b->stab.as<uint8_t>(), b->nullf.as<uint8_t>(),
                           b->dict.as<uint64_t>(), dt.view, dt.chi, dt.nrm, xd, d_part.as<double>());
        QMB_HIP(hipGetLastError());
        std::vector<double> part((size_t)blocks * 2);
        QMB_HIP(hipMemcpyAsync(part.data(), d_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        QMB_HIP(hipStreamSynchronize(ctx->stream));
        // <Psi|A|Psi> = conj(sum_r conj(x_r) (M x)_r) = sum_r x_r conj((M x)_r), summed in block order
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < blocks; ++i) {
            sr += part[2 * i];
            si += part[2 * i + 1];
        }
        out[0] = sr;
        out[1] = si;
    });
}

}  // extern "C"
