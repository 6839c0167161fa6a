// Host check of the term compiler (qmbcuda_amd/csrc/terms.hpp): for random
// operator lists (Sz, S+, S- on few sites, so repeated sites are common, zero
// scalars included) the compiled masks give the same applicability, image and
// scalar bits as applying the operators one by one, left to right (the
// reference's inner loop, Q/exact_diagonalization/utils.cu:133-161).
#include <cstdio>
#include <cstring>
#include <random>

#include "terms.hpp"

using namespace qmb;

static bool sequential(const qmb_terms &t, int h, uint64_t src, uint64_t &tgt, double &cr, double &ci) {
    tgt = src;
    cr = t.scalar[2 * h];
    ci = t.scalar[2 * h + 1];
    for (int k = 0; k < t.n_ops[h]; ++k) {
        const int s = t.site[h * t.max_ops + k], op = t.op_type[h * t.max_ops + k];
        const int spin = (int)((tgt >> s) & 1ULL);
        if (op == QMB_OP_SM) {
            if (!spin) return false;
            tgt &= ~(1ULL << s);
        } else if (op == QMB_OP_SP) {
            if (spin) return false;
            tgt |= 1ULL << s;
        } else {
            const double f = -1.0 + 2.0 * (double)spin;
            cr = cr * 0.5 * f;
            ci = ci * 0.5 * f;
        }
    }
    return true;
}

static bool compiled(const TermC &c, uint64_t src, uint64_t &tgt, double &cr, double &ci) {
    if ((src & c.req) != c.req_val) return false;
    tgt = (src & ~c.set) | c.set_val;
    const bool neg = (__builtin_popcountll(~src & c.sz) & 1) != 0;
    cr = neg ? -c.coef.x : c.coef.x;
    ci = neg ? -c.coef.y : c.coef.y;
    return true;
}

int main() {
    std::mt19937_64 rng(7);
    const int N = 12, MO = 5, NT = 3000;
    qmb_terms t{};
    t.n_sites = N;
    t.n_terms = NT;
    t.max_ops = MO;
    t.op_type.resize(NT * MO);
    t.site.resize(NT * MO);
    t.n_ops.resize(NT);
    t.scalar.resize(2 * NT);
    std::uniform_real_distribution<double> u(-2.0, 2.0);
    for (int h = 0; h < NT; ++h) {
        t.n_ops[h] = 1 + (int)(rng() % MO);
        for (int k = 0; k < MO; ++k) {
            t.op_type[h * MO + k] = (int)(rng() % 3);
            t.site[h * MO + k] = (int)(rng() % 4);
        }
        t.scalar[2 * h] = (h % 7 == 0) ? 0.0 : u(rng);
        t.scalar[2 * h + 1] = (h % 3 == 0) ? 0.0 : u(rng);
    }
    long acts = 0, bad = 0;
    for (int h = 0; h < NT; ++h) {
        const TermC c = compile_term(t, h);
        for (uint64_t x = 0; x < (1u << N); ++x) {
            uint64_t t1 = 0, t2 = 0;
            double a1 = 0, b1 = 0, a2 = 0, b2 = 0;
            const bool o = sequential(t, h, x, t1, a1, b1), n = compiled(c, x, t2, a2, b2);
            if (o != n) {
                ++bad;
                continue;
            }
            if (!o) continue;
            ++acts;
            if (t1 != t2 || std::memcmp(&a1, &a2, 8) || std::memcmp(&b1, &b2, 8)) ++bad;
        }
    }
    std::printf("terms %d applications %ld mismatches %ld\n", NT, acts, bad);
    return bad == 0 && acts > 0 ? 0 : 1;
}
