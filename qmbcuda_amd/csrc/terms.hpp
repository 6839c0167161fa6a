// Device form of a conjugated term table (the reference's BuildOperatorVecs
// output, Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336)
// and the application of one term to a basis state.
#pragma once

#include <algorithm>
#include <cmath>
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

// A term table uploaded to the device (int8 operators and sites, double2 scalars).
struct DeviceTerms {
    DevBuf op, site, n_ops, coef;
    TermsView view{};
    explicit DeviceTerms(const qmb_terms &t) {
        const int nt = t.n_terms, mo = t.max_ops;
        std::vector<int8_t> op8((size_t)std::max(1, nt * mo)), site8((size_t)std::max(1, nt * mo)),
            nops8((size_t)std::max(1, nt));
        std::vector<double> c((size_t)std::max(1, nt) * 2, 0.0);
        for (int i = 0; i < nt * mo; ++i) {
            op8[i] = (int8_t)t.op_type[i];
            site8[i] = (int8_t)(t.site[i] < 0 ? 0 : t.site[i]);
        }
        for (int h = 0; h < nt; ++h) {
            nops8[h] = (int8_t)t.n_ops[h];
            c[2 * h] = t.scalar[2 * h];
            c[2 * h + 1] = t.scalar[2 * h + 1];
        }
        op.alloc(op8.size());
        site.alloc(site8.size());
        n_ops.alloc(nops8.size());
        coef.alloc(c.size() * sizeof(double));
        QMB_HIP(hipMemcpy(op.p, op8.data(), op8.size(), hipMemcpyHostToDevice));
        QMB_HIP(hipMemcpy(site.p, site8.data(), site8.size(), hipMemcpyHostToDevice));
        QMB_HIP(hipMemcpy(n_ops.p, nops8.data(), nops8.size(), hipMemcpyHostToDevice));
        QMB_HIP(hipMemcpy(coef.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
        view = TermsView{nt, mo, op.as<int8_t>(), site.as<int8_t>(), coef.as<double2>(), n_ops.as<int8_t>()};
    }
};

// Orbit norms n(m) = sqrt(G m)/G for stabiliser sizes m = 0..G (entry 0 unused).
inline std::vector<double> orbit_norm_table(int G) {
    std::vector<double> nrm(G + 1, 0.0);
    for (int m = 1; m <= G; ++m) nrm[m] = std::sqrt((double)G * (double)m) / (double)G;
    return nrm;
}

}  // namespace qmb
