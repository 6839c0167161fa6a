/*
 * qmb_oracle.h -- FP64 CPU restatement of the reference's exact-diagonalisation
 * hot path (jeohjee/QMBCuda). TEST INFRASTRUCTURE ONLY: it is the checker that
 * the HIP library is compared against and the timed CPU baseline. Nothing in
 * the product library links or calls it.
 *
 * Conventions (all from the reference):
 *   - site s <-> bit s of the state, spin up = bit set
 *     (QMBCuda/exact_diagonalization/utils.cu:137,157)
 *   - group_el[g*LS + s] = destination bit of site s under element g
 *     (utils.cu:52-59, symmetries/SymmetryClass.cu:55-99, utils/misc_funcs.cu:80-90)
 *   - characters chars[(a*G + g)*2 + {0,1}] = (re, im) of chi_a(g)
 *     (symmetries/Groups.cu:89-112)
 *   - term tables are the conjugated tables of BuildOperatorVecs
 *     (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336):
 *     op types 0 = Sz, 1 = Sp, 2 = Sm (quantum_operators/SingleParticleOperators.h:4)
 *
 * The spin-flip extension (flip[g] = 1 -> image XOR all-ones mask) has no
 * reference counterpart; it is the same in the oracle and the HIP library.
 */
#ifndef QMB_ORACLE_H
#define QMB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t orc_binom(int n, int k);
int64_t orc_rank_ref(uint64_t state, int n_up);
int64_t orc_rank_colex(uint64_t state);
uint64_t orc_unrank_colex(int64_t rank, int n_up);
int64_t orc_sector_states(int n_sites, int n_up, uint64_t *out);
uint64_t orc_apply(uint64_t s, const uint32_t *perm_g, int flip, int n_sites);

void orc_rank_colex_many(const uint64_t *states, int64_t n, int64_t *ranks);
void orc_srs_min(int n_sites, int G, const uint32_t *perm, const uint8_t *flip,
                 const uint64_t *bas, int64_t n, int64_t *out_min);
int64_t orc_sort_unique(int64_t *v, int64_t n);
int orc_orbit_tables(int n_sites, int G, const uint32_t *perm, const uint8_t *flip,
                     int n_irreps, const double *chars, const uint64_t *bas, int64_t n_bas,
                     const int64_t *rep_rank, int64_t n_orb, int32_t *stab,
                     double *norm_sum, int64_t *orbit_idx);

int64_t orc_build_orbits(int n_sites, int n_up, int G, const uint32_t *perm,
                         const uint8_t *flip, int n_irreps, const double *chars,
                         const uint64_t *bas, int64_t n_bas, int64_t *rep_rank,
                         int32_t *stab, double *norm_sum, int64_t *orbit_idx);

void orc_closed_form_norms(int64_t n_orb, int G, int n_irreps, const int32_t *stab,
                           const double *norm_sum, double tol_norm, double *norms);

int orc_ell_build(int n_sites, int n_up, int G, int n_irreps, const double *chars,
                  int alpha, int n_terms, int max_ops, const int32_t *op_type,
                  const int32_t *site, const double *scalar, const int32_t *n_ops,
                  const uint64_t *bas, int64_t n_orb, const int64_t *rep_rank,
                  const double *norms, const int64_t *orbit_idx, double tol_norm,
                  int max_terms, int is_complex, uint32_t *ind, double *val,
                  int16_t *track);

int orc_ell_rows_from_reps(int n_sites, int G, const uint32_t *perm, const uint8_t *flip, int n_irreps,
                           const double *chars, int alpha, int n_terms, int max_ops, const int32_t *op_type,
                           const int32_t *site, const double *scalar, const int32_t *n_ops,
                           const uint64_t *reps, int64_t n_reps, const int64_t *rows, int64_t n_sample,
                           double tol_norm, int max_terms, int is_complex, uint32_t *ind, double *val,
                           int16_t *track);

void orc_expectation(int n_sites, int G, const uint32_t *group_el, const uint8_t *flip,
                     int n_irreps, const double *chars, int alpha, int n_terms, int max_ops,
                     const int32_t *op_type, const int32_t *site, const double *scalar,
                     const int32_t *n_ops, const uint64_t *bas, int64_t n_orb,
                     const int64_t *rep_rank, const double *norms, const int64_t *orbit_idx,
                     double tol_norm, int is_complex, const double *x, double *out);

void orc_spmv(int64_t n, int W, const uint32_t *ind, const double *val,
              const int16_t *track, int is_complex, const double *x, double *y);

void orc_start_vector(uint64_t seed, int64_t n, int is_complex, double *out);

int orc_lanczos(int64_t n, int W, const uint32_t *ind, const double *val,
                const int16_t *track, int is_complex, int n_iter, uint64_t seed,
                const double *start, double *alpha, double *beta);

int orc_lanczos_converge(int64_t n, int W, const uint32_t *ind, const double *val,
                         const int16_t *track, int is_complex, int max_iter,
                         int check_every, double tol, uint64_t seed,
                         const double *start, double *alpha, double *beta,
                         int *iters, double *e0);

double orc_lanczos_eigvec(int64_t n, int W, const uint32_t *ind, const double *val,
                          const int16_t *track, int is_complex, int n_iter,
                          uint64_t seed, const double *start, const double *krylov,
                          double e0, double *x);

int orc_tridiag_ql(int m, const double *d, const double *e, double *w, double *z);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
