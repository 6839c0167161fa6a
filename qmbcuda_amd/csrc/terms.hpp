// Device form of a conjugated term table (the reference's BuildOperatorVecs
// output, Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336)
// and the application of one term to a basis state.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace qmb {

struct TermsView {
    int n_terms, max_ops;
    const int8_t *op;      // [n_terms*max_ops]
    const int8_t *site;    // [n_terms*max_ops]
    const double2 *coef;   // [n_terms]
    const int8_t *n_ops;   // [n_terms]
};

// Conjugated operators of term h applied left to right to src (the reference's
// inner loop, utils.cu:133-161). Returns false when the term annihilates src.
template <class S>
__device__ inline bool apply_term(const TermsView &tv, int h, S src, S &tgt, double &cr, double &ci) {
    tgt = src;
    const double2 c0 = tv.coef[h];
    cr = c0.x;
    ci = c0.y;
    const int nk = tv.n_ops[h];
    for (int k = 0; k < nk; ++k) {
        const int s = tv.site[h * tv.max_ops + k];
        const int op = tv.op[h * tv.max_ops + k];
        const int spin = (int)((tgt >> s) & (S)1);
        if (op == QMB_OP_SM) {
            if (!spin) return false;
            tgt &= (S)~((S)1 << s);
        } else if (op == QMB_OP_SP) {
            if (spin) return false;
            tgt |= (S)1 << s;
        } else {
            const double f = -1.0 + 2.0 * (double)spin;
            cr = cr * 0.5 * f;
            ci = ci * 0.5 * f;
        }
    }
    return true;
}

// Orbit norms n(m) = sqrt(G m)/G for stabiliser sizes m = 0..G (entry 0 unused).
inline std::vector<double> orbit_norm_table(int G) {
    std::vector<double> nrm(G + 1, 0.0);
    for (int m = 1; m <= G; ++m) nrm[m] = std::sqrt((double)G * (double)m) / (double)G;
    return nrm;
}

// A term table uploaded to the device (int8 operators and sites, double2
// scalars) together with the irrep's character row and the orbit norm table of
// its group: one host blob, one copy.
struct DeviceTerms {
    DevBuf blob;
    TermsView view{};
    const double2 *chi = nullptr;  // [G]
    const double *nrm = nullptr;   // [G + 1]
    DeviceTerms(const qmb_terms &t, const double *chi_row, int G) {
        const int nt = t.n_terms, mo = t.max_ops;
        const size_t n_coef = (size_t)std::max(1, nt), n_ops8 = (size_t)std::max(1, nt * mo);
        auto up16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
        const size_t o_coef = 0, o_chi = up16(o_coef + n_coef * 16), o_nrm = up16(o_chi + (size_t)G * 16),
                     o_op = up16(o_nrm + (size_t)(G + 1) * 8), o_site = up16(o_op + n_ops8),
                     o_nops = up16(o_site + n_ops8), total = up16(o_nops + n_coef);
        std::vector<unsigned char> h(total, 0);
        double *coef = reinterpret_cast<double *>(h.data() + o_coef);
        for (int k = 0; k < nt; ++k) {
            coef[2 * k] = t.scalar[2 * k];
            coef[2 * k + 1] = t.scalar[2 * k + 1];
            h[o_nops + k] = (unsigned char)(int8_t)t.n_ops[k];
        }
        for (int i = 0; i < nt * mo; ++i) {
            h[o_op + i] = (unsigned char)(int8_t)t.op_type[i];
            h[o_site + i] = (unsigned char)(int8_t)(t.site[i] < 0 ? 0 : t.site[i]);
        }
        if (chi_row) std::memcpy(h.data() + o_chi, chi_row, (size_t)G * 16);
        const std::vector<double> nt_tab = orbit_norm_table(G);
        std::memcpy(h.data() + o_nrm, nt_tab.data(), nt_tab.size() * 8);
        blob.alloc(total);
        QMB_HIP(hipMemcpy(blob.p, h.data(), total, hipMemcpyHostToDevice));
        char *d = static_cast<char *>(blob.p);
        chi = reinterpret_cast<const double2 *>(d + o_chi);
        nrm = reinterpret_cast<const double *>(d + o_nrm);
        view = TermsView{nt,
                         mo,
                         reinterpret_cast<const int8_t *>(d + o_op),
                         reinterpret_cast<const int8_t *>(d + o_site),
                         reinterpret_cast<const double2 *>(d + o_coef),
                         reinterpret_cast<const int8_t *>(d + o_nops)};
    }
};

}  // namespace qmb
