// Device form of a conjugated term table (the reference's BuildOperatorVecs
// output, Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336)
// and the application of one term to a basis state.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.hpp"

namespace qmb {

// One term compiled on the host from its operator list (applied left to right,
// the reference's inner loop, utils.cu:133-161): the term acts on state x iff
// (x & req) == req_val; the image is (x & ~set) | set_val; its scalar is
// coef * 0.5^(number of Sz) times the signs of the Sz factors whose site value
// is known at that point (folded into coef), times -1 for every Sz site that
// still carries its original bit and has it 0 (the parity of ~x & sz). The
// factors are 1, +-0.5, so the products are exact in any order: the values
// equal the sequential ones bit for bit. A term that can never act has
// req = 1, req_val = 2. vclass: the term's coefficient class (terms whose coef
// has the same bit pattern share it), used to code ELL values.
struct TermC {
    uint64_t req, req_val, set, set_val, sz;
    double2 coef;
    uint64_t vclass;
};

struct TermsView {
    int n_terms;
    const TermC *t;  // [n_terms]
};

// The term h on src. Returns false when the term annihilates src.
template <class S>
__device__ inline bool apply_term(const TermsView &tv, int h, S src, S &tgt, double &cr, double &ci) {
    const TermC &t = tv.t[h];
    if ((src & (S)t.req) != (S)t.req_val) return false;
    tgt = (S)((src & (S)~(S)t.set) | (S)t.set_val);
    const bool neg = (popc_s((S)(~src & (S)t.sz)) & 1) != 0;
    cr = neg ? -t.coef.x : t.coef.x;
    ci = neg ? -t.coef.y : t.coef.y;
    return true;
}

// Host: the compiled form of term h of a table.
inline TermC compile_term(const qmb_terms &tt, int h) {
    TermC c{};
    uint64_t known = 0, kval = 0;  // sites whose current value is fixed, and that value
    double cr = tt.scalar[2 * h], ci = tt.scalar[2 * h + 1];
    bool never = false;
    for (int k = 0; k < tt.n_ops[h]; ++k) {
        const int s = tt.site[h * tt.max_ops + k];
        const int op = tt.op_type[h * tt.max_ops + k];
        const uint64_t b = 1ULL << s;
        if (op == QMB_OP_SM || op == QMB_OP_SP) {
            const uint64_t need = op == QMB_OP_SM ? b : 0;  // S- needs an up spin, S+ a down spin
            if (known & b) {
                if ((kval & b) != need) never = true;
            } else {
                c.req |= b;
                c.req_val |= need;
            }
            known |= b;
            kval = (kval & ~b) | (op == QMB_OP_SM ? 0 : b);
        } else {
            cr *= 0.5;
            ci *= 0.5;
            if (known & b) {
                if (!(kval & b)) {
                    cr = -cr;
                    ci = -ci;
                }
            } else {
                c.sz ^= b;  // sign from the original bit (applied twice: no sign)
            }
        }
    }
    if (never) {
        c.req = 1;
        c.req_val = 2;
    }
    c.set = known;
    c.set_val = kval;
    c.coef = make_double2(cr, ci);
    return c;
}

// Orbit norms n(m) = sqrt(G m)/G for stabiliser sizes m = 0..G (entry 0 unused).
inline std::vector<double> orbit_norm_table(int G) {
    std::vector<double> nrm(G + 1, 0.0);
    for (int m = 1; m <= G; ++m) nrm[m] = std::sqrt((double)G * (double)m) / (double)G;
    return nrm;
}

// Locality order of a term table: the terms sorted on (lowest site, then
// highest site) descending, stable. A term without operators counts as lying
// above every site. Basis rows are sorted by colex rank, which the highest
// occupied sites dominate, so the 64 rows of one chunk share their high sites:
// terms on high sites act on the same rows of the chunk, and when they come
// first, slot k of those rows comes from the same term and gathers nearby
// columns.
inline std::vector<int> locality_order(const qmb_terms &t) {
    std::vector<int> lo(t.n_terms), hi(t.n_terms), ord(t.n_terms);
    for (int h = 0; h < t.n_terms; ++h) {
        int a = t.n_sites, b = t.n_sites;
        if (t.n_ops[h] > 0) {
            a = t.n_sites;
            b = -1;
            for (int k = 0; k < t.n_ops[h]; ++k) {
                const int s = t.site[h * t.max_ops + k];
                a = std::min(a, s);
                b = std::max(b, s);
            }
        }
        lo[h] = a;
        hi[h] = b;
        ord[h] = h;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        if (lo[x] != lo[y]) return lo[x] > lo[y];
        return hi[x] > hi[y];
    });
    return ord;
}

// A term table uploaded to the device (compiled terms, in table order or in
// the given order) together with the irrep's character row and the orbit norm
// table of its group: one host blob, one copy.
struct DeviceTerms {
    DevBuf blob;
    TermsView view{};
    const double2 *chi = nullptr;  // [G]
    const double *nrm = nullptr;   // [G + 1]
    static constexpr size_t kMaxClasses = 1 << 14;
    std::vector<double2> classes;  // distinct term coefficients, in order of first use (TermC::vclass)
    DeviceTerms(const qmb_terms &t, const double *chi_row, int G, const std::vector<int> *order = nullptr) {
        const int nt = t.n_terms;
        auto up16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
        const size_t o_terms = 0, o_chi = up16((size_t)std::max(1, nt) * sizeof(TermC)),
                     o_nrm = up16(o_chi + (size_t)G * 16), total = up16(o_nrm + (size_t)(G + 1) * 8);
        std::vector<unsigned char> h(total, 0);
        TermC *tc = reinterpret_cast<TermC *>(h.data() + o_terms);
        for (int k = 0; k < nt; ++k) tc[k] = compile_term(t, order ? (*order)[k] : k);
        // classes of the terms that can leave the diagonal (the others never give
        // an off-diagonal entry: no operator changes a spin, or none acts); beyond
        // kMaxClasses the classes are not used and no longer counted
        for (int k = 0; k < nt; ++k) {
            tc[k].vclass = 0;
            if (tc[k].set == 0 || tc[k].req_val > tc[k].req || classes.size() > kMaxClasses) continue;
            for (const double2 &c : classes) {
                if (std::memcmp(&c, &tc[k].coef, sizeof(double2)) == 0) break;
                ++tc[k].vclass;
            }
            if (tc[k].vclass == classes.size()) classes.push_back(tc[k].coef);
        }
        if (chi_row) std::memcpy(h.data() + o_chi, chi_row, (size_t)G * 16);
        const std::vector<double> nt_tab = orbit_norm_table(G);
        std::memcpy(h.data() + o_nrm, nt_tab.data(), nt_tab.size() * 8);
        blob.alloc(total);
        QMB_HIP(hipMemcpy(blob.p, h.data(), total, hipMemcpyHostToDevice));
        char *d = static_cast<char *>(blob.p);
        chi = reinterpret_cast<const double2 *>(d + o_chi);
        nrm = reinterpret_cast<const double *>(d + o_nrm);
        view = TermsView{nt, reinterpret_cast<const TermC *>(d + o_terms)};
    }
};

}  // namespace qmb
