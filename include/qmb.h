/*
 * qmb.h -- C ABI of the MI355X exact-diagonalisation engine (libqmb.so).
 *
 * The reference (jeohjee/QMBCuda) has no FFI: its boundary is a C++ class API
 * plus one free function. Each entry point below names the reference code it
 * replaces (paths relative to the reference repository root, Q/ = QMBCuda/).
 * The C++ facade in include/qmb/ (the reference's class API) calls only these
 * functions.
 *
 * Rules of the ABI
 *  - plain pointers and sizes; no C++ types cross it; no exception escapes;
 *  - every function returns a qmb_status; qmb_last_error() gives the text of the
 *    last failure on the calling thread;
 *  - opaque handles own their device memory and are released by *_destroy;
 *    a handle holds the handles it was made from (a basis its context and
 *    group, an ELL its basis, a network its context and sample bases), so
 *    *_destroy calls may come in any order; host arrays belong to the caller;
 *  - a handle is used by one host thread at a time; every handle made from a
 *    context runs on that context's HIP stream, and a call may return while its
 *    last kernels still run there (every call that reads their results is
 *    ordered after them on the stream or synchronises it first);
 *  - states are uint64 (at most 64 sites); Lanczos runs in FP64.
 *
 * Conventions kept from the reference (bit-exact):
 *  - site s <-> bit s, spin up = bit set (Q/exact_diagonalization/utils.cu:137,157);
 *  - group_el[g*n_sites + s] = bit that site s moves to under element g
 *    (utils.cu:52-59; Q/symmetries/SymmetryClass.cu:55-99); element 0 is the identity;
 *  - chars[(a*G + g)*2 + {0,1}] = (re, im) of the character of irrep a at element g
 *    (Q/symmetries/Groups.cu:89-112);
 *  - operator types 0 = Sz, 1 = Sp, 2 = Sm (Q/quantum_operators/SingleParticleOperators.h:4);
 *  - term tables are the conjugated tables of BuildOperatorVecs
 *    (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336);
 *  - rows are orbit representatives (minimum integer of the orbit) in ascending order;
 *    ELL slot 0 is the diagonal; off-diagonal slots follow term order, unmerged
 *    (utils.cu:83-184).
 * New in this engine: uint64 states, an optional spin-flip bit per group element
 * (image XOR all-ones), the compact layout (only orbits compatible with the irrep).
 */
#ifndef QMB_H
#define QMB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    QMB_OK = 0,
    QMB_ERR_INVALID = 1,       /* bad argument or inconsistent input */
    QMB_ERR_HIP = 2,           /* a HIP runtime call failed */
    QMB_ERR_OOM = 3,           /* device or host allocation failed */
    QMB_ERR_NOT_CONVERGED = 4, /* Lanczos stopped at max_iter before its tolerance */
    QMB_ERR_NO_DEVICE = 5,     /* no usable GPU */
    QMB_ERR_UNSUPPORTED = 6    /* valid request this build does not implement */
} qmb_status;

enum { QMB_OP_SZ = 0, QMB_OP_SP = 1, QMB_OP_SM = 2 };
enum { QMB_LAYOUT_REFERENCE = 0, QMB_LAYOUT_COMPACT = 1 };
enum { QMB_REAL = 0, QMB_COMPLEX = 1 };

typedef struct qmb_ctx qmb_ctx;
typedef struct qmb_group qmb_group;
typedef struct qmb_terms qmb_terms;
typedef struct qmb_basis qmb_basis;
typedef struct qmb_ell qmb_ell;

/* ---- errors, version -------------------------------------------------------
 * Replaces CUDA_CHECK / CUBLAS_CHECK / CUSOLVER_CHECK exceptions
 * (Q/cublas_utils.h:45-62, Q/cusolver_utils.h:47-54). */
const char *qmb_last_error(void);
int qmb_version(void);

/* ---- context: device + one HIP stream ---------------------------------------
 * The reference runs everything on the default device and stream with a
 * cudaDeviceSynchronize after each kernel (Q/exact_diagonalization/lanczos.cu:83,214). */
qmb_status qmb_ctx_create(int device, qmb_ctx **out);
qmb_status qmb_ctx_destroy(qmb_ctx *ctx);
qmb_status qmb_ctx_synchronize(qmb_ctx *ctx);
/* Device count without creating a context (0 when no GPU). */
int qmb_device_count(void);

/* ---- symmetry group ---------------------------------------------------------
 * Replaces SymmetryClass<T>::GetGroupElemArr + the character table copy of the
 * driver (Q/symmetries/SymmetryClass.cu:26-71; Q/exact_diagonalization/
 * HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:45-61). flip may be NULL (no spin
 * flip). The element list must be closed under composition and G <= 128. */
qmb_status qmb_group_create(int n_sites, int n_elems, const uint32_t *group_el,
                            const uint8_t *flip, int n_irreps, const double *chars,
                            qmb_group **out);
qmb_status qmb_group_destroy(qmb_group *g);
/* valid[a] = 1 when character row a is a one-dimensional representation of the
 * group (chi(g h) = chi(g) chi(h)); qmb_group_create checks closure under
 * composition and marks the rows. Rows of a translation group combined with a
 * point operation are valid only at the momenta the operation fixes; bases of
 * the other rows are refused. */
qmb_status qmb_group_irreps_valid(const qmb_group *g, int32_t *valid);
/* Inverse-element table (inv[g] = index of g^-1) and 1 if the fast torus
 * (cyclic-translation) path was detected, 0 for the generic permutation path. */
qmb_status qmb_group_info(const qmb_group *g, int32_t *inv, int *torus_path);

/* ---- term table -------------------------------------------------------------
 * The flat, conjugated Hamiltonian tables of BuildOperatorVecs
 * (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:306-336): op_type/site are
 * [n_terms*max_ops] (site -1 pads), scalar [n_terms*2] (re, im), n_ops[n_terms].
 * Every term must conserve Sz (as many Sp as Sm), or QMB_ERR_INVALID. */
qmb_status qmb_terms_create(int n_sites, int n_terms, int max_ops, const int32_t *op_type,
                            const int32_t *site, const double *scalar,
                            const int32_t *n_ops, qmb_terms **out);
qmb_status qmb_terms_destroy(qmb_terms *t);

/* ---- symmetry-reduced basis (on device) -------------------------------------
 * Replaces BuildBasisStates + BuildSRStates (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:
 * 277-296, 341-400) and the kernels SetUpSRS, SolveSRS_min, SetUpSRS_from_indices,
 * ComputeSRSNorms, ComputeOrbitIndices (Q/exact_diagonalization/utils.cu:39-80,
 * 716-805). Rows = orbit representatives in ascending order. REFERENCE layout keeps
 * every orbit (null orbits of the irrep become empty rows, like the reference);
 * COMPACT keeps only the orbits compatible with the irrep. */
typedef struct {
    int64_t sector_dim; /* C(n_sites, n_up) */
    int64_t n_rows;     /* rows of the matrix */
    int64_t n_null;     /* rows that are null for the irrep (REFERENCE layout) */
    int32_t n_sites, n_up, irrep, layout;
    int64_t device_bytes;
} qmb_basis_info;

qmb_status qmb_basis_build(qmb_ctx *ctx, const qmb_group *g, int n_up, int irrep,
                           int layout, qmb_basis **out);
qmb_status qmb_basis_destroy(qmb_basis *b);
qmb_status qmb_basis_get_info(const qmb_basis *b, qmb_basis_info *info);
/* Host copies of the representatives, stabiliser sizes and null flags (any may be NULL). */
qmb_status qmb_basis_export(const qmb_basis *b, uint64_t *reps, uint16_t *stab, uint8_t *is_null);
/* Row lookup of arbitrary sector states (replaces the rank -> (orbit, g) table
 * orbit_indices, utils.cu:66-80): row[i] = row of the state's orbit or -1 if that
 * orbit is not a row; elem[i] = smallest g with image_g(rep) = state. */
qmb_status qmb_basis_lookup(const qmb_basis *b, int64_t n, const uint64_t *states,
                            int64_t *row, int32_t *elem);

/* ---- ELL Hamiltonian (on device) --------------------------------------------
 * Replaces BuildHamiltonianXXZForAbelianGroup (utils.cu:83-184) with its launch
 * (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:127-167). max_terms = the reference's
 * row width W; <= 0 means the reference rule n_terms*n_up/n_sites + 1 (:71-72).
 * A row that needs more than W slots is an error (the reference overflows). */
typedef struct {
    int64_t n_rows;
    int64_t nnz;          /* stored entries excluding padding (sum of row lengths) */
    int64_t stored;       /* entries including chunk padding */
    int32_t max_row_len;
    int32_t max_terms;    /* W of the reference layout */
    int32_t value_type;   /* QMB_REAL or QMB_COMPLEX */
    int32_t chunk_rows;   /* rows per SELL chunk (64 = one wavefront) */
    int64_t device_bytes;
    double spmv_bytes;    /* algorithmic bytes of one SpMV (int32 column + FP64 value per entry) */
    int32_t packed;       /* 1: packed values (below) */
    int32_t n_values;     /* distinct values in the table(s) of a packed ELL (complex: real + imaginary parts) */
    double format_bytes;  /* bytes of one SpMV in the stored format */
} qmb_ell_info;

/* A REAL matrix is stored packed when its distinct values fit a table of at most
 * 1024 entries and the index bits left above the column bits (32 - ceil(log2 n_rows)):
 * every off-diagonal slot is one 32-bit word col | (index << col_bits), the
 * diagonal an FP64 array. A COMPLEX matrix likewise when its distinct real parts
 * and its distinct imaginary parts each fit 1024 entries and both indexes fit
 * beside the column: col | (re index << col_bits) | (im index << (col_bits +
 * re bits)). Values and products are unchanged (SpMV results are bit-identical
 * to the unpacked form); the matrix stream is 4 instead of 12 (20 complex)
 * bytes per entry. value_type | QMB_KEEP_FP64_VALUES keeps the unpacked form.
 * While a packed matrix is built it takes 6 bytes of device memory per stored
 * entry (column word + 16-bit value code) when its term coefficients, character
 * classes and stabiliser sizes give at most 2^14 value codes, else the FP64 form's
 * 12 (20) bytes; QMB_VALUE_CODES=0 in the environment selects the second route
 * (same result). */
enum { QMB_KEEP_FP64_VALUES = 16 };
/* value_type | QMB_LOCALITY_SLOTS: the slots of each row follow the terms
 * sorted on (lowest site, then highest site) descending, stable (a term with no
 * operator first), instead of table order. Same matrix, same entries; each
 * row's sum runs in the new slot order. Rows are in colex order, so the 64 rows
 * of a chunk share their high sites and, with those terms first, slot k of the
 * chunk's rows gathers nearby columns (6x6 Sz=0 k=0 sector: 0.54 -> 0.34
 * distinct 128 B lines per gathered value, fused Lanczos SpMV 26.3 -> 24.8 ms).
 * No counterpart in the reference, whose slot order is the table order. */
enum { QMB_LOCALITY_SLOTS = 32 };
qmb_status qmb_ell_build(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *t,
                         int value_type, int max_terms, qmb_ell **out);
qmb_status qmb_ell_destroy(qmb_ell *e);
qmb_status qmb_ell_get_info(const qmb_ell *e, qmb_ell_info *info);
/* Reference row-major layout (ind[n_rows*W], val[n_rows*W] or [n_rows*W*2] for
 * complex, track[n_rows]), exactly as the reference kernel leaves its buffers. */
qmb_status qmb_ell_export_reference(const qmb_ell *e, uint32_t *ind, double *val, int16_t *track);
/* Rows [row0, row0 + n_rows) in the same layout (ind[n_rows*W], ...): a shard of
 * a matrix too large to export whole; column indices stay global. */
qmb_status qmb_ell_export_rows(const qmb_ell *e, int64_t row0, int64_t n_rows, uint32_t *ind, double *val,
                               int16_t *track);

/* ---- SpMV -------------------------------------------------------------------
 * Replaces SparseMatDenseVec_prod_EEL (utils.cu:601-624). Vectors are n_rows
 * doubles (REAL) or n_rows interleaved (re, im) pairs (COMPLEX). */
qmb_status qmb_spmv_host(qmb_ctx *ctx, const qmb_ell *e, const double *x, double *y);
qmb_status qmb_spmv_device(qmb_ctx *ctx, const qmb_ell *e, const double *x_dev, double *y_dev);
/* Runs reps back-to-back device SpMVs on internal vectors and returns the mean
 * kernel time in ms (HIP events on the context stream). */
qmb_status qmb_spmv_benchmark(qmb_ctx *ctx, const qmb_ell *e, int reps, double *ms_per_spmv);

/* ---- Lanczos ----------------------------------------------------------------
 * Replaces LanczosEigVals_CUDA (Q/exact_diagonalization/lanczos.cu:128-317) and
 * LanczosEigenStateFromSeed (:31-104) with an FP64 device loop: the same
 * three-term recurrence, a convergence test on the Ritz residual, optional
 * reorthogonalisation (needed for n_eigs > 1: classical Gram-Schmidt against
 * the kept vectors, repeated once when the first pass removed more than 1-1/sqrt(2)
 * of the norm), host FP64 tridiagonal solve. */
/* start vector: seeded random (device generated), user host array, user device array */
enum { QMB_START_RANDOM = 0, QMB_START_USER = 1, QMB_START_DEVICE = 2 };

typedef struct {
    int32_t max_iter;    /* upper bound on iterations */
    int32_t n_eigs;      /* lowest eigenvalues wanted (1..8) */
    int32_t check_every; /* iterations between convergence checks (0 -> 10) */
    int32_t reorth;      /* 0 plain recurrence, 1 full reorthogonalisation every step,
                          * 2 partial: only when the estimated loss of orthogonality
                          * passes sqrt(eps) (Simon 1984), 3 selective: against the Ritz
                          * vectors whose residual is below sqrt(eps) ||T|| (Parlett and
                          * Scott 1979; tested after every step, not with fixed_iter or
                          * checkpoints; single-sector solves only); 1-3 keep max_iter vectors */
    int32_t fixed_iter;  /* 1: run exactly max_iter steps, no test (reference semantics) */
    int32_t start;       /* QMB_START_RANDOM, QMB_START_USER or QMB_START_DEVICE */
    int32_t want_eigvec; /* 1: rebuild the ground-state vector (second pass) */
    int32_t profile;     /* k > 0: time every k-th SpMV (steps 0, k, 2k, ...) with HIP events */
    double tol;          /* eigenvalue tolerance, relative to max(1,|E|) (0 -> 1e-13) */
    uint64_t seed;       /* start-vector seed (reference default 2212456) */
    const double *start_vec; /* start vector: host (USER) or device (DEVICE) pointer, n_rows
                              * values (x2 interleaved re/im for a complex ELL) */
} qmb_lanczos_opts;

typedef struct {
    int32_t iters;       /* Lanczos steps taken */
    int32_t converged;   /* 1 if the tolerance was met */
    int32_t breakdown;   /* 1 if an invariant subspace was hit (exact result) */
    int32_t n_spmv;
    int32_t n_reorth;    /* steps that reorthogonalised */
    int32_t n_spmv_timed; /* SpMVs whose kernel time is in ms_spmv (profile) */
    double residual[8];  /* Ritz residuals of the returned eigenvalues */
    double eig_residual; /* ||H x - E x|| of the rebuilt vector (want_eigvec) */
    double ms_total;     /* wall time of the call */
    double ms_spmv;      /* summed kernel time of the timed SpMVs (profile) */
    double ms_host;      /* host tridiagonal work */
} qmb_lanczos_report;

void qmb_lanczos_default_opts(qmb_lanczos_opts *o);

/* evals[n_eigs]; krylov (optional, [max_iter]) = Krylov coefficients of the ground
 * state (lanczos.cu:268-279); eigvec (optional, host, n_rows values) = ground
 * state when want_eigvec. */
qmb_status qmb_lanczos(qmb_ctx *ctx, const qmb_ell *e, const qmb_lanczos_opts *opts,
                       double *evals, double *krylov, double *eigvec,
                       qmb_lanczos_report *report);

/* ---- on-disk formats: restartable runs (new) ---------------------------------
 * The reference has only a commented-out raw float dump of its ELL arrays
 * (Q/exact_diagonalization/HeisenbergHamAbelianSymms_CUDA.cu:149-162,
 * Q/utils/print_funcs.cu:73-90). qmb_ell_save writes the device matrix (FP64 or
 * packed) behind a versioned header tied to its basis (sector, irrep, layout,
 * a hash of the group tables); qmb_ell_load reads it back for the same basis
 * and checks on the device that every index stays inside the arrays.
 * qmb_lanczos_checkpointed is qmb_lanczos that writes its exact state to `path`
 * at the first convergence check at least `every` steps after the previous write
 * (0: only when stopping), stops after `stop_after` steps of this call (0: no
 * limit; returns QMB_ERR_NOT_CONVERGED with the current Ritz values), and with
 * resume = 1 continues from `path` instead of the start vector (same ELL and
 * options; a USER/DEVICE start vector must be passed again for want_eigvec).
 * A resumed run ends with the same results as an uninterrupted one. The state
 * holds a[], b[], two vectors, or every kept vector with reorthogonalisation. */
qmb_status qmb_ell_save(const qmb_ell *e, const char *path);
qmb_status qmb_ell_load(qmb_ctx *ctx, const qmb_basis *b, const char *path, qmb_ell **out);
qmb_status qmb_lanczos_checkpointed(qmb_ctx *ctx, const qmb_ell *e, const qmb_lanczos_opts *opts,
                                    const char *path, int every, int resume, int stop_after,
                                    double *evals, double *krylov, double *eigvec,
                                    qmb_lanczos_report *report);

/* Drop-in for the reference's free function LanczosEigVals_CUDA
 * (Q/exact_diagonalization/lanczos.h:21-32): fixed NIter steps on a
 * reference-layout ELL given as HOST arrays (index_mat[nobv*max_terms],
 * vals_mat[nobv*max_terms] (x2 if complex), track_vec[nobv]); E_tmp gets the
 * lowest eigenvalue; eig_vec_Krylov_out (optional) the Krylov coefficients; with
 * checkEigState the function value is ||Hx - Ex||/nobv like the reference's,
 * else 0. */
double qmb_lanczos_eigvals_ref(double *E_tmp, int NStates, int NIter, int nobv, int max_terms,
                               long long seed, const uint32_t *index_mat, const double *vals_mat,
                               const int16_t *track_vec, double *eig_vec_Krylov_out,
                               int exportKrylovVec, int checkEigState, int is_complex,
                               qmb_status *status);

/* The same drop-in with the reference's ownership: index_mat, vals_mat and
 * track_vec are DEVICE pointers of the default context's GPU (device 0), borrowed
 * like LanczosEigVals_CUDA borrows the driver's thrust::device_vector data
 * (Q/exact_diagonalization/lanczos.h:21-32; call site
 * HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:194-197). The matrix is converted to
 * the SELL-64 layout on the device (no copy through host memory); a track entry
 * outside [0, max_terms] or a column >= nobv is QMB_ERR_INVALID, and so is a
 * pointer that is not device memory. E_tmp and eig_vec_Krylov_out are host
 * pointers, as in the reference. */
double qmb_lanczos_eigvals_ref_device(double *E_tmp, int NStates, int NIter, int nobv, int max_terms,
                                      long long seed, const uint32_t *index_mat, const double *vals_mat,
                                      const int16_t *track_vec, double *eig_vec_Krylov_out,
                                      int exportKrylovVec, int checkEigState, int is_complex,
                                      qmb_status *status);

/* ---- expectation values ------------------------------------------------------
 * Replaces the device part of ComputeStaticExpValZeroT
 * (Q/exact_diagonalization/HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:421-601): the kernel
 * ExpecValArbOperatorAbelianIrrepProjHeisenbergZeroT (Q/exact_diagonalization/utils.cu:451-548)
 * and its reduction (:583-590). out[0..1] = <Psi|A|Psi> (re, im), where Psi is the
 * many-body state of the vector x of the basis's sector (n_rows values, interleaved
 * re/im when is_complex; x on the host, or on the device when x_on_device): for an
 * eigenvector of the sector matrix, its eigenstate (no normalisation is applied).
 * A is any Sz-conserving operator as a conjugated term table, in the same form as
 * the Hamiltonian's. It need not commute with the group: the sector sees its group
 * average (1/G) sum_g g A g^-1, taken over every image of each row's representative;
 * when A commutes with the group (checked on the host) the average is skipped.
 * flags = QMB_EXPECT_GROUP_AVERAGE forces the average; *averaged (optional) = 1 when
 * it was taken. Differences from the reference kernel: every image of the source
 * contributes (it stops at the first image a term annihilates), and the pairing of
 * x with its conjugate is that of <Psi|A|Psi> (the reference's equals it only for
 * vectors that are real up to a phase); FP64, no tol_norm. */
enum { QMB_EXPECT_GROUP_AVERAGE = 1 };
qmb_status qmb_expectation(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *a, int is_complex,
                           const double *x, int x_on_device, int flags, double *out, int *averaged);

/* qmb_expectation with nothing allocated: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). x_dev must be device
 * memory of the context's GPU holding the basis's n_rows values (interleaved
 * re/im when is_complex); the result is the same as qmb_expectation's, bit for
 * bit. The per-block partial sums, the compiled terms and the group tables live
 * in a 1 MiB scratch array that the library keeps on each GPU until the process
 * ends: a table of more than 15292 terms is refused (QMB_ERR_INVALID; use
 * qmb_expectation). Every call on one GPU shares that scratch: calls that
 * overlap on one GPU (from several threads or contexts) race. */
qmb_status qmb_expectation_noalloc(qmb_ctx *ctx, const qmb_basis *b, const qmb_terms *a, int is_complex,
                                   const double *x_dev, int flags, double *out, int *averaged);

/* ---- batched Lanczos over many sectors (new; no reference counterpart) -------
 * The reference solves irreps one after another (HeisenbergGenericXXZ_AbelianSymms_CUDA.cu:127-229).
 * This solves n independent ELLs together: one launch per phase covers every
 * instance (a work table of row chunks), the iteration chunk is captured once as
 * a HIP graph and replayed, converged instances drop out. evals[n*n_eigs],
 * reports[n] (may be NULL). reorth = 2 (partial) decides on the device, per
 * instance and step, inside the captured iteration. The host checks the
 * instances' tridiagonal problems on several threads while the next chunk runs.
 * opts->check_every is the number of iterations per graph replay (72 N=24
 * sectors: 5 is ~3 % faster than 10, 20 ~18 % slower). With opts->profile,
 * every report's ms_spmv is the summed GPU time of the graph replays (HIP
 * events) and n_spmv_timed their count; ms_total is the call's wall time and
 * ms_host the host's check time. */
qmb_status qmb_lanczos_batched(qmb_ctx *ctx, int n, const qmb_ell *const *ells,
                               const qmb_lanczos_opts *opts, double *evals,
                               qmb_lanczos_report *reports);
/* qmb_lanczos_batched with every instance's ground-state vector: vecs[i] (host,
 * n_rows doubles, 2 n_rows interleaved for a complex ELL) gets x = sum_j z_j
 * phi_j, z the unit eigenvector of the instance's tridiagonal matrix for its
 * lowest eigenvalue (z_0 >= 0), phi_j its kept Lanczos vectors - the vector
 * qmb_lanczos returns with want_eigvec and reorthogonalisation. Needs
 * opts->reorth = 1 or 2 (the vectors are kept); opts->want_eigvec is ignored.
 * For many small sectors at once (training ground states of a learned start).
 * No counterpart in the reference (its eigenvector pass, lanczos.cu:31-104,
 * serves one irrep at a time). */
qmb_status qmb_lanczos_batched_vectors(qmb_ctx *ctx, int n, const qmb_ell *const *ells,
                                       const qmb_lanczos_opts *opts, double *evals, double *const *vecs,
                                       qmb_lanczos_report *reports);

/* ---- learned start vector (new; no reference counterpart) -------------------
 * A translation-invariant network over 8-spin windows (+-1 spins, Delta, J2/J1
 * -> tanh 128 -> tanh 128 -> log-amplitude, phase), log psi(s) = sum over the
 * windows of the lattice (the N windows of an N-site ring, or the windows set
 * by qmb_mlp_set_windows), plus i*pi*(up spins on sublattice A) (the Marshall
 * sign rule as a fixed prior: the network learns the phase relative to it).
 * Bf16 weights and activations on MFMA, FP32 accumulation, FP32 Adam master
 * weights. Parameter layout: W1[128][10], b1[128],
 * W2[128][128], b2[128], W3[2][128], b3[2]. Samples are fully symmetric bases
 * (k = 0, flip-even; compact layout) with their ED ground states; the loss is the
 * mean infidelity 1 - |<u, c>|^2/(<c,c><u,u>), c_r = psi(r) sqrt(G/m_r). */
typedef struct qmb_mlp qmb_mlp;

int qmb_mlp_num_params(void);
qmb_status qmb_mlp_create(qmb_ctx *ctx, uint64_t seed, qmb_mlp **out);
qmb_status qmb_mlp_destroy(qmb_mlp *m);
qmb_status qmb_mlp_get_params(const qmb_mlp *m, float *theta);
qmb_status qmb_mlp_set_params(qmb_mlp *m, const float *theta);
/* Network outputs for all 256 window patterns of n couplings: out[(i*256 + p)*2 + {a, phi}]. */
qmb_status qmb_mlp_forward(qmb_mlp *m, int n, const double *delta, const double *j2, float *out);
/* Windows of a basis that is not a plain ring (e.g. a 2-D torus): n_win windows
 * of 8 site indices (sites[w*8 + k] = site of bit k of window w), and the
 * sublattice-A mask of the Marshall sign prior. A plain ring (translation
 * group G = N) defaults to the N windows q..q+7 (mod N), sublattice A = even
 * sites. The basis must be the fully symmetric sector (k = 0, flip-even),
 * compact layout; call before adding its samples. */
qmb_status qmb_mlp_set_windows(qmb_mlp *m, const qmb_basis *b, int n_win, const uint8_t *sites,
                               uint64_t sublattice_a);
/* The network keeps the basis alive while the sample is held. */
qmb_status qmb_mlp_add_sample(qmb_mlp *m, const qmb_basis *b, const double *ground_state, double delta, double j2);
qmb_status qmb_mlp_clear_samples(qmb_mlp *m);
/* Mean loss over the samples and (grad != NULL) its gradient, host FP32 [num_params]. */
qmb_status qmb_mlp_loss_grad(qmb_mlp *m, double *loss, float *grad);
/* Full-batch Adam steps; loss_history[steps] optional (loss before each step). */
qmb_status qmb_mlp_train(qmb_mlp *m, int steps, double lr, double *loss_history);
/* Real start vector x_r = |c_r| cos(beta_r - beta_ref) on a fully symmetric basis
 * (host, n_rows). *fallback = 1 when it is not finite or zero: use the random start. */
qmb_status qmb_mlp_start_vector(qmb_mlp *m, const qmb_basis *b, double delta, double j2, double *x,
                                int *fallback);
/* Lanczos on e (built on basis b) from the learned start vector, which stays on
 * the device: the network, the start vector and the Lanczos run share one
 * stream and only two norms are read back. The start is x/||x|| + 0.1 xi/||xi||
 * with xi_r = 2u_r - 1 from the SplitMix64 stream of opts->seed (the noise keeps
 * every eigenvector in reach when x misses the ground state). opts may be NULL
 * (defaults); its start fields are overridden. The Lanczos residual judges the
 * learned start, and the seeded random start is the fallback: *fallback
 * (optional) = 0 when the learned run is the result; 1 when the learned vector
 * was unusable (zero or not finite) and only the random-start run was made; 2
 * when the learned run's Ritz residual missed the tolerance within max_iter
 * (QMB_ERR_NOT_CONVERGED) and the random-start run, with the same options,
 * is the result. eigvec as in qmb_lanczos. */
qmb_status qmb_lanczos_learned(qmb_mlp *m, const qmb_basis *b, const qmb_ell *e, double delta, double j2,
                               const qmb_lanczos_opts *opts, double *evals, double *eigvec,
                               qmb_lanczos_report *report, int *fallback);

#ifdef __cplusplus
}
#endif
#endif
