/*
 * qmb_oracle.c -- FP64 CPU restatement of the reference's ED hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library and
 * the CPU baseline that the benchmark times beside it; the product library
 * never links, loads or calls it. Each function names the reference code it
 * restates (paths relative to the reference repository root, Q/ =
 * QMBCuda/).
 *
 * Parity pins (what fixes each function independently of this file):
 *   orc_binom            exact integer Pascal table; tests compare with math.comb
 *   orc_rank_ref         literal restatement of the reference rank formula; tests
 *                        check it equals orc_rank_colex and the position in the
 *                        ascending enumeration, N <= 20, all states
 *   orc_sector_states    closed form: ascending popcount-k integers, count C(N,k)
 *   orc_build_orbits     representative count = Burnside count; norms equal the
 *                        closed form sqrt(G*m)/G; brute force on tiny N
 *   orc_ell_build        spectral completeness: the union of all sector spectra
 *                        equals the dense full-sector spectrum (numpy, N <= 14);
 *                        Hermiticity after merging duplicate columns
 *   orc_ell_rows_from_reps  equal, row for row and bit for bit, to orc_ell_build on
 *                        every orbit of small sectors (chains with flip, 4x4, all
 *                        irreps); a representative removed from the list is
 *                        reported missing
 *   orc_spmv             dense matvec of the exported matrix (numpy)
 *   orc_expectation      <Psi|A|Psi> computed densely in the full Sz sector from
 *                        the operator algebra (numpy), with Psi embedded through
 *                        orbit states that are checked to carry H's eigenvectors
 *                        (H Psi = E Psi); covers complex irreps, spin flips,
 *                        operators that do not commute with the group and
 *                        non-Hermitian operators
 *   orc_lanczos*         dense eigenvalues (numpy) and closed forms
 *                        (Majumdar-Ghosh, 2-site, 4-ring, known E0 values)
 *   orc_tridiag_ql       numpy.linalg.eigh on random symmetric tridiagonals
 *   orc_start_vector     SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): tests
 *                        compare with an independent Python SplitMix64
 */
#include "qmb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Exact binomial coefficient from Pascal's rule. The reference uses the
 * floating recursion NChoosek2 (Q/utils/misc_funcs.cu:23-32) rounded with +0.5
 * at its call sites; for n <= 64 the integer table is the exact value it
 * approximates. */
uint64_t orc_binom(int n, int k) {
    static uint64_t table[65][65];
    static int ready = 0;
    if (!ready) {
        for (int a = 0; a <= 64; a++) {
            table[a][0] = 1;
            for (int b = 1; b <= 64; b++)
                table[a][b] = (a == 0) ? 0 : table[a - 1][b - 1] + table[a - 1][b];
        }
        ready = 1;
    }
    if (n < 0 || k < 0 || n > 64 || k > 64) return 0;
    return table[n][k];
}

/* Floating binomial recursion of the reference, restated
 * (Q/utils/misc_funcs_gpu.cu:9-18, identical host copy Q/utils/misc_funcs.cu:23-32). */
static double ref_nchoosek2(double num, double den) {
    if (den <= 0) return 1;
    if (num <= 0) return 1;
    if (den <= 1) return num;
    return (num / den) * ref_nchoosek2(num - 1, den - 1);
}

/* Literal restatement of GetIndexInHilbertSubspace (Q/exact_diagonalization/
 * utils.cu:9-36): walk the bits from the top, and for every set bit add the
 * rounded floating binomial. The reference walks 32 bits; this walks 64 so it
 * also covers the wider states of the new build (the formula is unchanged). */
int64_t orc_rank_ref(uint64_t state, int n_up) {
    int64_t acc = 0;
    int curr = 0;
    int tot = 64;
    for (int bit = 63; bit >= 0; bit--) {
        if ((state >> bit) & 1ULL) {
            int delta_x = tot - (n_up - curr);
            int x_coord = tot;
            curr += 1;
            if (delta_x > 0) {
                double b = ref_nchoosek2((double)(x_coord - 1), (double)(delta_x - 1));
                acc += (int64_t)(b + 0.5);
            }
        }
        tot -= 1;
    }
    return acc;
}

/* Colexicographic rank: sum over the set bits p_0 < p_1 < ... of C(p_j, j+1).
 * Equal to orc_rank_ref (pinned by test) and to the position of the state in
 * the ascending popcount-k enumeration. */
int64_t orc_rank_colex(uint64_t state) {
    int64_t r = 0;
    int j = 0;
    while (state) {
        int p = __builtin_ctzll(state);
        r += (int64_t)orc_binom(p, j + 1);
        j++;
        state &= state - 1;
    }
    return r;
}

uint64_t orc_unrank_colex(int64_t rank, int n_up) {
    uint64_t s = 0;
    uint64_t r = (uint64_t)rank;
    for (int j = n_up; j >= 1; j--) {
        int p = j - 1;
        while (orc_binom(p + 1, j) <= r) p++;
        s |= 1ULL << p;
        r -= orc_binom(p, j);
    }
    return s;
}

/* BuildBasisStates (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:277-296):
 * loop over every integer below 2^LS and keep those with popcount == GS_sector,
 * in ascending order. Returns the count (or -1 if out is too small to be
 * meaningful: the caller sizes it as C(N,k)). */
int64_t orc_sector_states(int n_sites, int n_up, uint64_t *out) {
    if (n_sites < 1 || n_sites > 40 || n_up < 0 || n_up > n_sites) return -1;
    uint64_t lim = 1ULL << n_sites;
    int64_t c = 0;
    for (uint64_t i = 0; i < lim; i++)
        if (__builtin_popcountll(i) == n_up) out[c++] = i;
    return c;
}

/* Bit permutation of SetUpSRS (Q/exact_diagonalization/utils.cu:49-59): every
 * set bit at site s moves to bit group_el[g][s]. The optional flip (new in this
 * build) complements all LS bits afterwards. */
uint64_t orc_apply(uint64_t s, const uint32_t *perm_g, int flip, int n_sites) {
    uint64_t f = 0;
    for (int site = n_sites - 1; site >= 0; site--)
        if ((s >> site) & 1ULL) f += 1ULL << perm_g[site];
    if (flip) {
        uint64_t mask = (n_sites == 64) ? ~0ULL : ((1ULL << n_sites) - 1ULL);
        f ^= mask;
    }
    return f;
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* Colex ranks of many states (GetIndexInHilbertSubspace, utils.cu:9-36, per state). */
void orc_rank_colex_many(const uint64_t *states, int64_t n, int64_t *ranks) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) ranks[i] = orc_rank_colex(states[i]);
}

/* SetUpSRS + SolveSRS_min (Q/exact_diagonalization/utils.cu:39-64, 750-762):
 * out_min[i] = min over g of rank(image of bas[i] under g). The reference keeps
 * the whole C(N,k)*G rank table in device memory first; the minimum is the
 * same. */
void orc_srs_min(int n_sites, int G, const uint32_t *perm, const uint8_t *flip,
                 const uint64_t *bas, int64_t n, int64_t *out_min) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t mn = -1;
        for (int g = 0; g < G; g++) {
            uint64_t img = orc_apply(bas[i], perm + (size_t)g * n_sites, flip ? flip[g] : 0, n_sites);
            int64_t r = orc_rank_colex(img);
            if (mn < 0 || r < mn) mn = r;
        }
        out_min[i] = mn;
    }
}

/* thrust::sort + unique_count + unique (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:367-370). */
int64_t orc_sort_unique(int64_t *v, int64_t n) {
    qsort(v, (size_t)n, sizeof(int64_t), cmp_i64);
    int64_t u = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || v[i] != v[i - 1]) v[u++] = v[i];
    return u;
}

/* SetUpSRS_from_indices + ComputeSRSNorms + ComputeOrbitIndices on the sorted
 * representative ranks (see orc_build_orbits). orbit_idx must hold 2*n_bas. */
int orc_orbit_tables(int n_sites, int G, const uint32_t *perm, const uint8_t *flip,
                     int n_irreps, const double *chars, const uint64_t *bas, int64_t n_bas,
                     const int64_t *rep_rank, int64_t n_orb, int32_t *stab,
                     double *norm_sum, int64_t *orbit_idx) {
    int64_t *srs = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_orb > 0 ? n_orb : 1) * (size_t)G);
    if (!srs) return -1;
#pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < n_orb; o++) {
        for (int g = 0; g < G; g++) {
            uint64_t img = orc_apply(bas[rep_rank[o]], perm + (size_t)g * n_sites, flip ? flip[g] : 0, n_sites);
            srs[o * G + g] = orc_rank_colex(img);
        }
        int m = 0;
        for (int g = 0; g < G; g++) m += (srs[o * G + g] == rep_rank[o]);
        stab[o] = m;
        for (int a = 0; a < n_irreps; a++) {
            double sr = 0.0, si = 0.0;
            const double *ch = chars + (size_t)a * G * 2;
            for (int i = 0; i < G; i++) {
                for (int j = 0; j < G; j++) {
                    if (bas[srs[o * G + i]] != bas[srs[o * G + j]]) continue;
                    double ar = ch[2 * i], ai = -ch[2 * i + 1];
                    double br = ch[2 * j], bi = ch[2 * j + 1];
                    sr += ar * br - ai * bi;
                    si += ar * bi + ai * br;
                }
            }
            (void)si;
            norm_sum[o * n_irreps + a] = sqrt(fabs(sr)) / (double)G;
        }
    }
    for (int64_t i = 0; i < 2 * n_bas; i++) orbit_idx[i] = -1;
    for (int64_t o = 0; o < n_orb; o++) {
        for (int g = 0; g < G; g++) {
            int64_t t = srs[o * G + g];
            if (orbit_idx[2 * t] < 0) {
                orbit_idx[2 * t] = o;
                orbit_idx[2 * t + 1] = g;
            }
        }
    }
    free(srs);
    return 0;
}

/* Reference-form orbit construction, restating BuildSRStates
 * (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:341-400):
 *   SetUpSRS (utils.cu:39-64)        SRS[i*G+g] = rank(image of state i under g)
 *   SolveSRS_min (utils.cu:750-762)  representative rank = min over g
 *   thrust sort/unique (…cu:367-370) sorted unique representatives
 *   SetUpSRS_from_indices (utils.cu:716-745)  orbit table of each representative
 *   ComputeSRSNorms (utils.cu:767-805)        norm[o*NIr+a] = sqrt|sum_{i,j: s_i=s_j}
 *                                             conj(chi_a(i)) chi_a(j)| / G
 *   ComputeOrbitIndices (utils.cu:66-80)      rank -> (orbit, g)
 * The reference's ComputeOrbitIndices lets all G threads of an orbit race; the
 * restatement keeps the smallest g (deterministic; any g of the coset gives the
 * same matrix element on compatible orbits).
 * Also returns the stabiliser size m of every orbit (the number of g whose image
 * is the representative itself). Returns n_orb, or -1 on bad input. */
int64_t orc_build_orbits(int n_sites, int n_up, int G, const uint32_t *perm,
                         const uint8_t *flip, int n_irreps, const double *chars,
                         const uint64_t *bas, int64_t n_bas, int64_t *rep_rank,
                         int32_t *stab, double *norm_sum, int64_t *orbit_idx) {
    if (flip) {
        for (int g = 0; g < G; g++)
            if (flip[g] && 2 * n_up != n_sites) return -1;
    }
    int64_t *srs_min = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_bas > 0 ? n_bas : 1));
    if (!srs_min) return -1;
    orc_srs_min(n_sites, G, perm, flip, bas, n_bas, srs_min);
    int64_t n_orb = orc_sort_unique(srs_min, n_bas);
    memcpy(rep_rank, srs_min, sizeof(int64_t) * (size_t)n_orb);
    free(srs_min);
    if (orc_orbit_tables(n_sites, G, perm, flip, n_irreps, chars, bas, n_bas, rep_rank, n_orb,
                         stab, norm_sum, orbit_idx))
        return -1;
    return n_orb;
}

/* Closed form of the reference norm: sqrt(G*m)/G when the character is trivial
 * on the stabiliser (the double sum is then G*m), else 0. The compatibility is
 * read off the double-sum norms: a null orbit has norm_sum < tol_norm, the
 * reference's own test (utils.cu:119,172). */
void orc_closed_form_norms(int64_t n_orb, int G, int n_irreps, const int32_t *stab,
                           const double *norm_sum, double tol_norm, double *norms) {
    for (int64_t o = 0; o < n_orb; o++)
        for (int a = 0; a < n_irreps; a++) {
            double ns = norm_sum[o * n_irreps + a];
            norms[o * n_irreps + a] =
                (ns < tol_norm) ? 0.0 : sqrt((double)G * (double)stab[o]) / (double)G;
        }
}

/* BuildHamiltonianXXZForAbelianGroup (Q/exact_diagonalization/utils.cu:83-184),
 * one row per orbit, reference layout (row-major, W = max_terms slots):
 *   slot 0: column = the row itself, value = accumulated diagonal;
 *   null row (norm < tol): slot 0 column written, track stays 0;
 *   then, in term order, conjugated operators applied left to right; an
 *   annihilated term is skipped; target == source adds to the diagonal; else
 *   the target's (orbit, g) is looked up and, unless that orbit is null, a new
 *   slot gets column = orbit and value = (coef * chi_alpha(g)) * (norm_o/norm_row).
 * Complex products are written out as (ar*br - ai*bi, ar*bi + ai*br) so the
 * HIP kernel can reproduce every bit. Returns 0, or 1 if a row needs more
 * than W slots (the reference would write past the row). */
typedef struct {
    int n_irreps, alpha, n_terms, max_ops, W, is_complex;
    const double *ch; /* characters of irrep alpha, (re, im) per element */
    const int32_t *op_type, *site, *n_ops;
    const double *scalar;
    double tol_norm;
} row_spec;

/* Target state -> its orbit's row o, the element gm with image_gm(rep_o) = target,
 * and the norm of that orbit for the irrep. Returns 1 to use the entry, 0 to skip
 * it (null orbit), a negative code on an inconsistency. */
typedef int (*target_fn)(void *lk, uint64_t tgt, int64_t *o, int64_t *gm, double *on);

/* One row of BuildHamiltonianXXZForAbelianGroup (utils.cu:115-181): slot 0 is
 * the diagonal (column = the row), then the off-diagonal targets in term order.
 * ind/val/track point at the row. Returns 0, 1 on overflow of W, or the lookup's
 * negative code. */
static int ell_row(const row_spec *r, uint64_t src, double cur, int64_t id, target_fn look, void *lk,
                   uint32_t *ind, double *val, int16_t *track) {
    const int W = r->W;
    ind[0] = (uint32_t)id;
    int slot = 1;
    if (cur < r->tol_norm) return 0;
    double dr = 0.0, di = 0.0;
    for (int hi = 0; hi < r->n_terms; hi++) {
        uint64_t tgt = src;
        double cr = r->scalar[2 * hi], ci = r->scalar[2 * hi + 1];
        int dead = 0;
        for (int k = 0; k < r->n_ops[hi]; k++) {
            int s = r->site[hi * r->max_ops + k];
            int op = r->op_type[hi * r->max_ops + k];
            int spin = (int)((tgt >> s) & 1ULL);
            if (op == 2) { /* Sm */
                if (!spin) { dead = 1; break; }
                tgt &= ~(1ULL << s);
            } else if (op == 1) { /* Sp */
                if (spin) { dead = 1; break; }
                tgt |= 1ULL << s;
            } else if (op == 0) { /* Sz */
                double f = -1.0 + 2.0 * (double)spin;
                cr = cr * 0.5 * f;
                ci = ci * 0.5 * f;
            } else {
                dead = 1;
                break;
            }
        }
        if (dead) continue;
        if (tgt == src) {
            dr += cr;
            di += ci;
            continue;
        }
        int64_t o, gm;
        double on;
        int use = look(lk, tgt, &o, &gm, &on);
        if (use < 0) return use;
        if (!use) continue;
        if (slot >= W) return 1;
        double ratio = on / cur;
        double xr = r->ch[2 * gm], xi = r->ch[2 * gm + 1];
        double pr = cr * xr - ci * xi;
        double pi = cr * xi + ci * xr;
        ind[slot] = (uint32_t)o;
        if (r->is_complex) {
            val[2 * slot] = pr * ratio;
            val[2 * slot + 1] = pi * ratio;
        } else {
            val[slot] = pr * ratio;
        }
        slot++;
    }
    if (r->is_complex) {
        val[0] = dr;
        val[1] = di;
    } else {
        val[0] = dr;
    }
    *track = (int16_t)slot;
    return 0;
}

typedef struct {
    const int64_t *orbit_idx;
    const double *norms;
    int n_irreps, alpha;
    double tol_norm;
} table_lookup;

/* ComputeOrbitIndices table (utils.cu:66-80): rank -> (orbit, g). */
static int lookup_table(void *lk, uint64_t tgt, int64_t *o, int64_t *gm, double *on) {
    const table_lookup *t = (const table_lookup *)lk;
    int64_t rk = orc_rank_colex(tgt);
    *o = t->orbit_idx[2 * rk];
    *gm = t->orbit_idx[2 * rk + 1];
    *on = t->norms[*o * t->n_irreps + t->alpha];
    return *on >= t->tol_norm;
}

int orc_ell_build(int n_sites, int n_up, int G, int n_irreps, const double *chars,
                  int alpha, int n_terms, int max_ops, const int32_t *op_type,
                  const int32_t *site, const double *scalar, const int32_t *n_ops,
                  const uint64_t *bas, int64_t n_orb, const int64_t *rep_rank,
                  const double *norms, const int64_t *orbit_idx, double tol_norm,
                  int max_terms, int is_complex, uint32_t *ind, double *val,
                  int16_t *track) {
    (void)n_sites;
    (void)n_up;
    const int W = max_terms;
    const int vs = is_complex ? 2 : 1;
    row_spec r = {n_irreps, alpha, n_terms, max_ops, W, is_complex, chars + (size_t)alpha * G * 2,
                  op_type, site, n_ops, scalar, tol_norm};
    table_lookup lk = {orbit_idx, norms, n_irreps, alpha, tol_norm};
    int overflow = 0;
    memset(ind, 0, sizeof(uint32_t) * (size_t)n_orb * W);
    memset(val, 0, sizeof(double) * (size_t)n_orb * W * vs);
    memset(track, 0, sizeof(int16_t) * (size_t)n_orb);
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
    for (int64_t id = 0; id < n_orb; id++) {
        int rc = ell_row(&r, bas[rep_rank[id]], norms[id * n_irreps + alpha], id, lookup_table, &lk,
                         ind + (size_t)W * id, val + (size_t)W * id * vs, track + id);
        overflow |= (rc != 0);
    }
    return overflow;
}

/* Orbit data of one state by brute force over the G images (SetUpSRS +
 * SolveSRS_min + SetUpSRS_from_indices + ComputeSRSNorms for a single orbit,
 * utils.cu:39-64,716-805): the orbit minimum, its stabiliser size, the
 * closed-form norm (0 when the double-sum norm is below tol_norm), and the
 * smallest element mapping the minimum onto the state. imgs holds G words. */
typedef struct {
    int n_sites, G, n_irreps, alpha;
    const uint32_t *perm;
    const uint8_t *flip;
    const double *ch;
    double tol_norm;
    const uint64_t *reps; /* ascending representatives of the rows */
    int64_t n_reps;
} brute_lookup;

static uint64_t brute_orbit(const brute_lookup *b, uint64_t s, uint64_t *imgs, int *stab, double *norm,
                            int64_t *g_to_s) {
    const int G = b->G;
    uint64_t mn = ~0ULL;
    for (int g = 0; g < G; g++) {
        uint64_t im = orc_apply(s, b->perm + (size_t)g * b->n_sites, b->flip ? b->flip[g] : 0, b->n_sites);
        if (im < mn) mn = im;
    }
    for (int g = 0; g < G; g++)
        imgs[g] = orc_apply(mn, b->perm + (size_t)g * b->n_sites, b->flip ? b->flip[g] : 0, b->n_sites);
    int m = 0;
    *g_to_s = -1;
    for (int g = 0; g < G; g++) {
        m += (imgs[g] == mn);
        if (*g_to_s < 0 && imgs[g] == s) *g_to_s = g;
    }
    double sr = 0.0;
    for (int i = 0; i < G; i++)
        for (int j = 0; j < G; j++) {
            if (imgs[i] != imgs[j]) continue;
            double ar = b->ch[2 * i], ai = -b->ch[2 * i + 1];
            double br = b->ch[2 * j], bi = b->ch[2 * j + 1];
            sr += ar * br - ai * bi;
        }
    double ns = sqrt(fabs(sr)) / (double)G;
    *stab = m;
    *norm = (ns < b->tol_norm) ? 0.0 : sqrt((double)G * (double)m) / (double)G;
    return mn;
}

static int64_t find_rep(const brute_lookup *b, uint64_t v) {
    int64_t lo = 0, hi = b->n_reps;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (b->reps[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return (lo < b->n_reps && b->reps[lo] == v) ? lo : -1;
}

static int lookup_brute(void *lk, uint64_t tgt, int64_t *o, int64_t *gm, double *on) {
    const brute_lookup *b = (const brute_lookup *)lk;
    uint64_t imgs[256];
    int m;
    uint64_t mn = brute_orbit(b, tgt, imgs, &m, on, gm);
    *o = find_rep(b, mn);
    if (*o < 0) return (*on >= b->tol_norm) ? -3 : 0; /* a compatible orbit missing from the list */
    return *on >= b->tol_norm;
}

/* Rows of the sector matrix for given row indices into an ascending list of
 * representatives (a list too large for the full reference tables, e.g. a
 * 36-site sector): each row is built as in orc_ell_build, with the orbit data of
 * the row and of every target found by brute force over the G images and the
 * target's row by binary search in the list (rows of null orbits are not in a
 * compact list and are skipped like the reference's null targets). Returns 0,
 * 1 on overflow of W, 2 if a listed representative is not its orbit's minimum,
 * 3 if a target's compatible orbit is missing from the list. */
int orc_ell_rows_from_reps(int n_sites, int G, const uint32_t *perm, const uint8_t *flip, int n_irreps,
                           const double *chars, int alpha, int n_terms, int max_ops, const int32_t *op_type,
                           const int32_t *site, const double *scalar, const int32_t *n_ops,
                           const uint64_t *reps, int64_t n_reps, const int64_t *rows, int64_t n_sample,
                           double tol_norm, int max_terms, int is_complex, uint32_t *ind, double *val,
                           int16_t *track) {
    if (G > 256) return -1;
    const int W = max_terms;
    const int vs = is_complex ? 2 : 1;
    row_spec r = {n_irreps, alpha, n_terms, max_ops, W, is_complex, chars + (size_t)alpha * G * 2,
                  op_type, site, n_ops, scalar, tol_norm};
    brute_lookup b = {n_sites, G, n_irreps, alpha, perm, flip, chars + (size_t)alpha * G * 2, tol_norm,
                      reps, n_reps};
    int err = 0;
    memset(ind, 0, sizeof(uint32_t) * (size_t)n_sample * W);
    memset(val, 0, sizeof(double) * (size_t)n_sample * W * vs);
    memset(track, 0, sizeof(int16_t) * (size_t)n_sample);
#pragma omp parallel for schedule(dynamic, 16) reduction(max : err)
    for (int64_t i = 0; i < n_sample; i++) {
        int64_t id = rows[i];
        uint64_t src = reps[id];
        uint64_t imgs[256];
        int m;
        double cur;
        int64_t g0;
        if (brute_orbit(&b, src, imgs, &m, &cur, &g0) != src) {
            err = err > 2 ? err : 2;
            continue;
        }
        int rc = ell_row(&r, src, cur, id, lookup_brute, &b, ind + (size_t)W * i, val + (size_t)W * i * vs,
                         track + i);
        int e = rc < 0 ? -rc : rc;
        err = err > e ? err : e;
    }
    return err;
}

/* <Psi|A|Psi> for the sector vector x (reference layout rows), the restatement of
 * ExpecValArbOperatorAbelianIrrepProjHeisenbergZeroT (Q/exact_diagonalization/
 * utils.cu:451-548) and its reduction and division by G
 * (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:583-590):
 * for every source row, every image g.rep (gi = 0..G-1, the reference's
 * SRS_states row) and every term of the conjugated table A: apply the term,
 * find the target's orbit o and element gm (orbit_idx), skip null orbits and
 * accumulate c * conj(chi(gi)) * chi(gm) * (n_o / n_src) * x_o into acc; then
 * out = sum_src x_src * conj(acc) / G. Two corrections to the reference: the
 * image loop continues past an image the term annihilates (the reference breaks
 * out of it), and x is paired as in <Psi|A|Psi> with Psi = sum conj(x_r) v_r
 * (the reference's conj(phi_src) ... phi_o differs for complex vectors). */
void orc_expectation(int n_sites, int G, const uint32_t *group_el, const uint8_t *flip,
                     int n_irreps, const double *chars, int alpha, int n_terms, int max_ops,
                     const int32_t *op_type, const int32_t *site, const double *scalar,
                     const int32_t *n_ops, const uint64_t *bas, int64_t n_orb,
                     const int64_t *rep_rank, const double *norms, const int64_t *orbit_idx,
                     double tol_norm, int is_complex, const double *x, double *out) {
    const double *ch = chars + (size_t)alpha * G * 2;
    double tr = 0.0, ti = 0.0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tr, ti)
    for (int64_t id = 0; id < n_orb; id++) {
        double cur = norms[id * n_irreps + alpha];
        if (cur < tol_norm) continue;
        uint64_t rep = bas[rep_rank[id]];
        double ar = 0.0, ai = 0.0;
        for (int gi = 0; gi < G; gi++) {
            uint64_t img = orc_apply(rep, group_el + (size_t)gi * n_sites, flip ? flip[gi] : 0, n_sites);
            double jr = ch[2 * gi], ji = -ch[2 * gi + 1]; /* conj(chi(gi)) */
            for (int hi = 0; hi < n_terms; hi++) {
                uint64_t tgt = img;
                double cr = scalar[2 * hi], ci = scalar[2 * hi + 1];
                int dead = 0;
                for (int k = 0; k < n_ops[hi]; k++) {
                    int s = site[hi * max_ops + k];
                    int op = op_type[hi * max_ops + k];
                    int spin = (int)((tgt >> s) & 1ULL);
                    if (op == 2) {
                        if (!spin) { dead = 1; break; }
                        tgt &= ~(1ULL << s);
                    } else if (op == 1) {
                        if (spin) { dead = 1; break; }
                        tgt |= 1ULL << s;
                    } else {
                        double f = -1.0 + 2.0 * (double)spin;
                        cr = cr * 0.5 * f;
                        ci = ci * 0.5 * f;
                    }
                }
                if (dead) continue;
                int64_t t = orc_rank_colex(tgt);
                int64_t o = orbit_idx[2 * t];
                int64_t gm = orbit_idx[2 * t + 1];
                double on = norms[o * n_irreps + alpha];
                if (on < tol_norm) continue;
                double ratio = on / cur;
                double qr = jr * ch[2 * gm] - ji * ch[2 * gm + 1];
                double qi = jr * ch[2 * gm + 1] + ji * ch[2 * gm];
                double fr = (cr * qr - ci * qi) * ratio, fi = (cr * qi + ci * qr) * ratio;
                double xr = is_complex ? x[2 * o] : x[o], xi = is_complex ? x[2 * o + 1] : 0.0;
                ar += fr * xr - fi * xi;
                ai += fr * xi + fi * xr;
            }
        }
        double sr = is_complex ? x[2 * id] : x[id], si = is_complex ? x[2 * id + 1] : 0.0;
        tr += (sr * ar + si * ai) / (double)G;
        ti += (si * ar - sr * ai) / (double)G;
    }
    out[0] = tr;
    out[1] = ti;
}

/* SparseMatDenseVec_prod_EEL (Q/exact_diagonalization/utils.cu:601-624):
 * y[r] = sum_{i < track[r]} val[r*W+i] * x[ind[r*W+i]]. */
void orc_spmv(int64_t n, int W, const uint32_t *ind, const double *val,
              const int16_t *track, int is_complex, const double *x, double *y) {
    if (!is_complex) {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; r++) {
            double acc = 0.0;
            for (int i = 0; i < track[r]; i++) acc += val[(size_t)r * W + i] * x[ind[(size_t)r * W + i]];
            y[r] = acc;
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; r++) {
            double ar = 0.0, ai = 0.0;
            for (int i = 0; i < track[r]; i++) {
                size_t w = (size_t)r * W + i;
                double vr = val[2 * w], vi = val[2 * w + 1];
                double xr = x[2 * (size_t)ind[w]], xi = x[2 * (size_t)ind[w] + 1];
                ar += vr * xr - vi * xi;
                ai += vr * xi + vi * xr;
            }
            y[2 * r] = ar;
            y[2 * r + 1] = ai;
        }
    }
}

/* Seeded uniform (0,1] start vector. The reference draws from cuRAND XORWOW
 * (Q/utils/misc_funcs_gpu.cu:31-52); that generator is not reproduced. This
 * uses the SplitMix64 finaliser (Steele, Lea, Flood 2014) as a counter-based
 * generator: draw c = mix(seed + (c+1)*0x9E3779B97F4A7C15); u = ((z>>11)+1)*2^-53.
 * Real vectors use counter i for entry i; complex ones use 2i (re), 2i+1 (im). */
static uint64_t splitmix_draw(uint64_t seed, uint64_t c) {
    uint64_t z = seed + (c + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_start_vector(uint64_t seed, int64_t n, int is_complex, double *out) {
    int64_t cnt = is_complex ? 2 * n : n;
    for (int64_t i = 0; i < cnt; i++)
        out[i] = (double)((splitmix_draw(seed, (uint64_t)i) >> 11) + 1ULL) * (1.0 / 9007199254740992.0);
}

static double vec_dot_re(int64_t n, int cplx, const double *a, const double *b) {
    double s = 0.0;
    int64_t cnt = cplx ? 2 * n : n;
    for (int64_t i = 0; i < cnt; i++) s += a[i] * b[i];
    return s; /* Re<a,b> with conj on a: re*re + im*im */
}

/* One Lanczos recurrence step engine shared by the fixed-length and the
 * converging drivers; restates LanczosEigVals_CUDA (Q/exact_diagonalization/
 * lanczos.cu:158-235) in FP64:
 *   phi normalised start; b[0] = 0;
 *   for n: b[n] = previous norm (n > 0); y = H phi; a[n] = Re<phi, y>;
 *          y += (-b_prev) * phi_prev; phi_prev = phi; phi = (-a[n]) * phi + y;
 *          norm = ||phi||; phi *= 1/norm.
 * beta[m] (the norm after the last step) is also returned: it is the Ritz
 * residual factor. */
typedef struct {
    int64_t n;
    int W, cplx;
    const uint32_t *ind;
    const double *val;
    const int16_t *track;
    double *phi, *prev, *y;
    double nrm;
} lz_state;

static int lz_init(lz_state *s, uint64_t seed, const double *start) {
    int64_t cnt = s->cplx ? 2 * s->n : s->n;
    s->phi = (double *)malloc(sizeof(double) * cnt);
    s->prev = (double *)calloc((size_t)cnt, sizeof(double));
    s->y = (double *)malloc(sizeof(double) * cnt);
    if (!s->phi || !s->prev || !s->y) return -1;
    if (start)
        memcpy(s->phi, start, sizeof(double) * cnt);
    else
        orc_start_vector(seed, s->n, s->cplx, s->phi);
    double nr = sqrt(vec_dot_re(s->n, s->cplx, s->phi, s->phi));
    if (!(nr > 0.0)) return -1;
    double inv = 1.0 / nr;
    for (int64_t i = 0; i < cnt; i++) s->phi[i] *= inv;
    s->nrm = 0.0;
    return 0;
}

static void lz_free(lz_state *s) {
    free(s->phi);
    free(s->prev);
    free(s->y);
}

/* Returns 0, or 1 on breakdown (the new norm is zero: invariant subspace). */
static int lz_step(lz_state *s, double *a_out, double *b_new) {
    int64_t cnt = s->cplx ? 2 * s->n : s->n;
    orc_spmv(s->n, s->W, s->ind, s->val, s->track, s->cplx, s->phi, s->y);
    double a = vec_dot_re(s->n, s->cplx, s->phi, s->y);
    double bp = s->nrm;
    for (int64_t i = 0; i < cnt; i++) {
        double yi = s->y[i] + (-bp) * s->prev[i];
        s->prev[i] = s->phi[i];
        s->phi[i] = (-a) * s->phi[i] + yi;
    }
    double nr = sqrt(vec_dot_re(s->n, s->cplx, s->phi, s->phi));
    *a_out = a;
    *b_new = nr;
    s->nrm = nr;
    if (!(nr > 0.0)) return 1;
    double inv = 1.0 / nr;
    for (int64_t i = 0; i < cnt; i++) s->phi[i] *= inv;
    return 0;
}

int orc_lanczos(int64_t n, int W, const uint32_t *ind, const double *val,
                const int16_t *track, int is_complex, int n_iter, uint64_t seed,
                const double *start, double *alpha, double *beta) {
    lz_state s = {n, W, is_complex, ind, val, track, NULL, NULL, NULL, 0.0};
    if (lz_init(&s, seed, start)) {
        lz_free(&s);
        return -1;
    }
    beta[0] = 0.0;
    int it = 0;
    for (; it < n_iter; it++) {
        double a, b;
        int brk = lz_step(&s, &a, &b);
        alpha[it] = a;
        beta[it + 1] = b;
        if (brk) {
            it++;
            break;
        }
    }
    lz_free(&s);
    return it;
}

/* Lowest Ritz value and residual of T_m (diag a[0..m-1], off-diagonal b[1..m-1]);
 * residual = b[m] * |last component of the lowest eigenvector|. */
static int ritz_lowest(int m, const double *a, const double *b, double *e0, double *res, double *gap) {
    double *d = (double *)malloc(sizeof(double) * m);
    double *e = (double *)malloc(sizeof(double) * (m > 1 ? m - 1 : 1));
    double *w = (double *)malloc(sizeof(double) * m);
    double *z = (double *)malloc(sizeof(double) * (size_t)m * m);
    for (int i = 0; i < m; i++) d[i] = a[i];
    for (int i = 0; i + 1 < m; i++) e[i] = b[i + 1];
    int rc = orc_tridiag_ql(m, d, e, w, z);
    if (rc == 0) {
        *e0 = w[0];
        *res = fabs(b[m] * z[(size_t)(m - 1)]); /* column 0, row m-1 */
        *gap = (m > 1) ? (w[1] - w[0]) : 1.0;
    }
    free(d);
    free(e);
    free(w);
    free(z);
    return rc;
}

/* Converging driver: the same recurrence, checked every check_every steps.
 * Converged when the lowest Ritz residual r satisfies r <= tol^(1/2) * max(1,|E|)
 * and r*r / max(gap, tol^(1/2)) <= tol * max(1,|E|) (Kato-Temple style bound on
 * the eigenvalue error), or on breakdown (exact). Returns 0 converged,
 * 1 not converged within max_iter, -1 error. */
int orc_lanczos_converge(int64_t n, int W, const uint32_t *ind, const double *val,
                         const int16_t *track, int is_complex, int max_iter,
                         int check_every, double tol, uint64_t seed,
                         const double *start, double *alpha, double *beta,
                         int *iters, double *e0) {
    lz_state s = {n, W, is_complex, ind, val, track, NULL, NULL, NULL, 0.0};
    if (lz_init(&s, seed, start)) {
        lz_free(&s);
        return -1;
    }
    if (check_every < 1) check_every = 1;
    beta[0] = 0.0;
    int status = 1;
    int it = 0;
    double sq = sqrt(tol);
    while (it < max_iter) {
        double a, b;
        int brk = lz_step(&s, &a, &b);
        alpha[it] = a;
        beta[it + 1] = b;
        it++;
        if (brk || it % check_every == 0 || it == max_iter) {
            double e, r, gp;
            if (ritz_lowest(it, alpha, beta, &e, &r, &gp)) {
                status = -1;
                break;
            }
            *e0 = e;
            double sc = fabs(e) > 1.0 ? fabs(e) : 1.0;
            if (brk || (r <= sq * sc && r * r / (gp > sq ? gp : sq) <= tol * sc)) {
                status = 0;
                break;
            }
        }
    }
    *iters = it;
    lz_free(&s);
    return status;
}

/* LanczosEigenStateFromSeed (Q/exact_diagonalization/lanczos.cu:31-104): replay
 * the recurrence from the same start and accumulate x = sum_n krylov[n] phi_n;
 * then return ||H x - e0 x|| (lanczos.cu:281-305 divides it by nobv; the caller
 * does that if it wants the reference's number). */
double orc_lanczos_eigvec(int64_t n, int W, const uint32_t *ind, const double *val,
                          const int16_t *track, int is_complex, int n_iter,
                          uint64_t seed, const double *start, const double *krylov,
                          double e0, double *x) {
    lz_state s = {n, W, is_complex, ind, val, track, NULL, NULL, NULL, 0.0};
    int64_t cnt = is_complex ? 2 * n : n;
    if (lz_init(&s, seed, start)) {
        lz_free(&s);
        return -1.0;
    }
    memset(x, 0, sizeof(double) * cnt);
    for (int it = 0; it < n_iter; it++) {
        for (int64_t i = 0; i < cnt; i++) x[i] = x[i] + krylov[it] * s.phi[i];
        double a, b;
        if (lz_step(&s, &a, &b)) break;
    }
    orc_spmv(n, W, ind, val, track, is_complex, x, s.y);
    double acc = 0.0;
    for (int64_t i = 0; i < cnt; i++) {
        double r = s.y[i] - e0 * x[i];
        acc += r * r;
    }
    lz_free(&s);
    return sqrt(acc);
}

/* Symmetric tridiagonal eigensolver: implicit QL with Wilkinson-type shifts
 * (Bowdler, Martin, Reinsch, Wilkinson, Numer. Math. 11 (1968) 293; Golub & Van
 * Loan, Matrix Computations, sec. 8.3). d[0..m-1] diagonal, e[0..m-2]
 * off-diagonal. On return w holds the eigenvalues in ascending order and, if z
 * is non-NULL, z[i + m*j] is component i of eigenvector j. Returns 0, or 1 if
 * an eigenvalue needed more than 60 sweeps. Stands in for the reference's FP32
 * cusolverDnXsyevd call (Q/cuda_wrappers/CudaWrappers.cu:174-266). */
int orc_tridiag_ql(int m, const double *d_in, const double *e_in, double *w, double *z) {
    double *d = w;
    double *e = (double *)calloc((size_t)m + 1, sizeof(double));
    for (int i = 0; i < m; i++) d[i] = d_in[i];
    for (int i = 0; i + 1 < m; i++) e[i] = e_in[i];
    if (z) {
        memset(z, 0, sizeof(double) * (size_t)m * m);
        for (int i = 0; i < m; i++) z[i + (size_t)m * i] = 1.0;
    }
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < m; l++) {
        int sweeps = 0;
        for (;;) {
            int mm = l;
            while (mm < m - 1) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= eps * dd) break;
                mm++;
            }
            if (mm == l) break;
            if (++sweeps > 60) {
                free(e);
                return 1;
            }
            /* shift from the leading 2x2 block of the unreduced part */
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            int underflow = 0;
            for (i = mm - 1; i >= l; i--) {
                double f = s * e[i];
                double b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    underflow = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    double *zi = z + (size_t)m * i;
                    double *zi1 = z + (size_t)m * (i + 1);
                    for (int k = 0; k < m; k++) {
                        double t = zi1[k];
                        zi1[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
            }
            if (underflow) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    free(e);
    /* selection sort ascending, carrying vectors */
    for (int i = 0; i < m - 1; i++) {
        int k = i;
        for (int j = i + 1; j < m; j++)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i];
            d[i] = d[k];
            d[k] = t;
            if (z) {
                for (int q = 0; q < m; q++) {
                    double u = z[q + (size_t)m * i];
                    z[q + (size_t)m * i] = z[q + (size_t)m * k];
                    z[q + (size_t)m * k] = u;
                }
            }
        }
    }
    return 0;
}
