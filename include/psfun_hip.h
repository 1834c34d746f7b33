/*
 * psfun_hip.h -- C-ABI of the MI355X (gfx950) f(A)x library.
 *
 * This is the drop-in boundary under PSFUN's Fortran API.  Every entry point
 * takes plain pointers and sizes, returns an int status (PSFUN_SUCCESS == 0)
 * and never stops or aborts the process.  Vector arguments are DEVICE
 * pointers (FP64, or interleaved (re,im) FP64 pairs for complex data) unless
 * the name says "host".  Work is ordered on the context's HIP stream; calls
 * that return a scalar to the host synchronise that stream.
 *
 * Reference interfaces replaced (paths relative to the PSFUN source tree):
 *   psb_spmm / psb_gedot / psb_genrm2 / psb_geamax / psb_geaxpby / v%scal
 *       -- PSBLAS primitives called from krylov/polynomial/psfun_d_lanczos.f90:183-377,
 *          krylov/polynomial/psfun_d_arnoldi.f90:183-329, krylov/shifted/psfun_dcg.F90:259-326
 *   psfun_d_lanczos  -- krylov/polynomial/psfun_d_lanczos.f90:37-400
 *   psfun_d_arnoldi  -- krylov/polynomial/psfun_d_arnoldi.f90:37-352
 *   psfun_d_saiarnoldi -- krylov/rational/psfun_d_sai.f90:37-419
 *   psfun_dcg_vect   -- krylov/shifted/psfun_dcg.F90:129-373
 *   psfun_d_serial%apply (f(H) step, EXP family) -- serial/psfun_d_serial_mod.F90:172-523
 *   hhtmethod1       -- quadratures/psfun_z_quadrules_mod.f90:40-85
 *   ellipkkp, ellipj -- utils/psfun_d_utils_mod.f90:37-99, utils/psfun_z_utils_mod.f90:39-107
 */
#ifndef PSFUN_HIP_H
#define PSFUN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* The Fortran layer maps these onto the PSBLAS error names used by the
 * reference (psb_success_, psb_err_invalid_istop_, ...). */
enum {
  PSFUN_SUCCESS = 0,
  PSFUN_ERR_INVALID_ARG = 1,          /* bad size / null pointer / bad option value */
  PSFUN_ERR_INVALID_ISTOP = 2,        /* psb_err_invalid_istop_ (psfun_d_lanczos.f90:108-113) */
  PSFUN_ERR_INVALID_VECT_STATE = 3,   /* psb_err_invalid_vect_state_ (psfun_d_lanczos.f90:133-142) */
  PSFUN_ERR_INVALID_ARGS_COMBINATION = 4, /* psb_err_invalid_args_combination_ */
  PSFUN_ERR_NOT_SQUARE = 5,           /* psfun_d_serial_mod.F90:206-210 */
  PSFUN_ERR_UNSUPPORTED = 6,          /* function/variant without a device implementation */
  PSFUN_ERR_ALLOC = 7,                /* device or pinned allocation failed */
  PSFUN_ERR_HIP = 8,                  /* a HIP runtime call failed */
  PSFUN_ERR_NO_DEVICE = 9,
  PSFUN_ERR_BREAKDOWN = 10,           /* solver breakdown reported as an error (CG/COCG only) */
  PSFUN_ERR_TOO_LARGE = 11            /* input exceeds the fixed scratch of psfun_lanczos_apply_noalloc */
};

/* f(H) step selection: function names follow psfun_d_serial%fname ('EXP',
 * 'USERF'), variants follow psfun_d_serial%variant
 * (serial/psfun_d_serial_mod.F90:214-498).  EXP takes the first six variants;
 * USERF takes SYM only. */
enum { PSFUN_FN_EXP = 0, PSFUN_FN_USERF = 1 };
enum {
  PSFUN_VAR_CHBSYM = 0,   /* symmetric H (Lanczos) */
  PSFUN_VAR_CHBHES = 1,   /* Hessenberg H (Arnoldi) */
  PSFUN_VAR_CHBGEN = 2,
  PSFUN_VAR_GENPADE = 3,
  PSFUN_VAR_TAYLOR = 4,
  PSFUN_VAR_SASQ = 5,
  PSFUN_VAR_SYM = 6       /* USERF: y = Q diag(f(scaling*lambda)) Q^T x of the
                             symmetric matrix read from H's upper triangle */
};

/* User function of the USERF f(H) step: f(x), called on the host with the
 * `user` pointer given to psfun_set_userf (psfun_d_serial%f_ptr,
 * serial/psfun_d_serial_mod.F90:49, 71-80; the reference's optional
 * derivative order is never passed by its SYM variant). */
typedef double (*psfun_userf_t)(double x, void *user);

typedef struct psfun_ctx *psfun_ctx_t;
typedef struct psfun_csr *psfun_csr_t;

/* Memory space of pointers handed to psfun_csr_create. */
enum { PSFUN_MEM_HOST = 0, PSFUN_MEM_DEVICE = 1 };

/* Storage the SpMV kernels read (psfun_csr_set_format). */
enum { PSFUN_FORMAT_AUTO = 0, PSFUN_FORMAT_CSR = 1, PSFUN_FORMAT_SELL = 2, PSFUN_FORMAT_HDIA = 3,
       PSFUN_FORMAT_CSRS = 4 };

/* ---- context / errors --------------------------------------------------- */
const char *psfun_strerror(int status);
/* Create a context on HIP device `device`.  The context owns a default-priority
 * stream, a pinned host staging area and a grow-only device workspace.
 * Every call reuses them (and the loop flags the host paces on), so one
 * context, with the matrices and batches created on it, serves one host
 * thread at a time; concurrent callers use one context each (contexts on one
 * device are independent). */
int psfun_ctx_create(int device, psfun_ctx_t *ctx);
/* Matrices and batches created on a context may still be destroyed after it
 * (they keep their device, not the context); any other use of them after
 * their context is gone is an error. */
int psfun_ctx_destroy(psfun_ctx_t ctx);
/* Order all work on an external hipStream_t (e.g. the caller's current
 * stream); NULL selects the HIP null stream. */
int psfun_ctx_set_stream(psfun_ctx_t ctx, void *hip_stream);
/* Go back to the context's own stream. */
int psfun_ctx_use_own_stream(psfun_ctx_t ctx);
/* The user function that fname = PSFUN_FN_USERF evaluates (fun%f_ptr set by
 * psfun_d_setpointer, serial/psfun_d_serial_mod.F90:153-168).  It is called
 * from the calling thread inside the apply, once per eigenvalue of each f(H)
 * step; USERF applies without a function return PSFUN_ERR_INVALID_ARG. */
int psfun_set_userf(psfun_ctx_t ctx, psfun_userf_t f, void *user);
/* The degree of the diagonal Pade approximant of the GENPADE variant
 * (fun%padedegree, set by psfun_d_serial%set('PADE_DEGREE', v) and passed to
 * DGPADM at serial/psfun_d_serial_mod.F90:271; default 6), for the serial
 * apply and the Arnoldi and shift-and-invert f(H) steps of this context.
 * Degree 6 evaluates exp(H) to the unit roundoff by the Taylor scheme (as
 * DGPADM(6) does, to roundoff); any other degree evaluates DGPADM's rational
 * approximant with its scaling and squaring, whose truncation error the
 * result then carries.  2 <= degree <= 30. */
int psfun_set_pade_degree(psfun_ctx_t ctx, int degree);
int psfun_ctx_synchronize(psfun_ctx_t ctx);
/* Number of compute units of the context's device (for launch-shape tuning). */
int psfun_ctx_num_cus(psfun_ctx_t ctx, int *num_cus);

/* ---- per-kernel timing (HIP events around each launch of a hot kernel) --- */
/* Kernel ids for psfun_prof_read. */
enum {
  PSFUN_PROF_LZ_SPMV = 0,    /* Lanczos SpMV + alpha dot (fused)          */
  PSFUN_PROF_LZ_UPDATE = 1,  /* Lanczos 3-term update + ||w||^2 partials */
  PSFUN_PROF_LZ_ASSEMBLE = 2,/* y = ||x|| V yk                           */
  PSFUN_PROF_SPMV = 3,       /* plain CSR SpMV (psfun_spmv)              */
  PSFUN_PROF_MS_SPMV = 4,    /* multi-shift CG seed SpMV + dots          */
  PSFUN_PROF_MS_UPDATE = 5,  /* multi-shift CG per-shift updates         */
  PSFUN_PROF_CG = 6,         /* shifted CG fused kernels                 */
  PSFUN_PROF_BATCH = 7,      /* batched Lanczos step kernels             */
  PSFUN_PROF_SDDMM = 8,      /* gradient SDDMM                           */
  PSFUN_PROF_AR_SPMV = 9,    /* Arnoldi SpMV step (+ control block)      */
  PSFUN_PROF_AR_MGS = 10,    /* Arnoldi Gram-Schmidt of one column       */
  PSFUN_PROF_REDUCE = 11,    /* dot / nrm2 / amax partial sums           */
  PSFUN_PROF_AXPBY = 12      /* axpby                                    */
};
/* Turn timing on/off (off by default); reset clears the totals.  on = 1
 * brackets each timed launch by two events; on = R in 2..16 also launches the
 * fused Lanczos SpMV step R times back to back inside its event pair (the
 * step is idempotent, results are unchanged), which amortises the events'
 * own cost in the per-launch average. */
int psfun_prof_enable(psfun_ctx_t ctx, int on);
int psfun_prof_reset(psfun_ctx_t ctx);
/* Totals since the last reset: milliseconds, launches, algorithmic bytes. */
int psfun_prof_read(psfun_ctx_t ctx, int kernel_id, double *total_ms, int64_t *launches, double *bytes);

/* ---- CSR matrix handle (psb_dspmat_type, CSR storage) ------------------- */
/* Host arrays (PSFUN_MEM_HOST) are copied to the device and converted to
 * 0-based indices; device arrays (PSFUN_MEM_DEVICE) must be 0-based and are
 * borrowed (the caller keeps them alive).  index_base is 0 or 1 (PSBLAS CSR is
 * 1-based).  Square matrices only. */
int psfun_csr_create(psfun_ctx_t ctx, int64_t n, int64_t nnz, const int32_t *rowptr,
                     const int32_t *colind, const double *val, int index_base, int mem,
                     psfun_csr_t *A);
int psfun_csr_destroy(psfun_csr_t A);
/* Replace the nonzero values (same pattern); `val` is a device pointer. */
int psfun_csr_set_values(psfun_csr_t A, const double *val);
/* Force the SpMV lanes-per-row (0 = automatic; 1..64 power of two). */
int psfun_csr_set_lanes(psfun_csr_t A, int lanes_per_row);
int psfun_csr_info(psfun_csr_t A, int64_t *n, int64_t *nnz, int *lanes_per_row);
/* Choose the storage the SpMV kernels read: PSFUN_FORMAT_CSR (vector CSR,
 * lanes_per_row lanes per row), PSFUN_FORMAT_SELL (sliced ELL, 64-row slices,
 * one thread per row; built at creation), PSFUN_FORMAT_HDIA (diagonal slices,
 * PSBLAS's "hacked DIA": each 64-row slice stores one value per row for each
 * distinct offset col - row in it, so the SpMV reads no column indices; built
 * on first choice, PSFUN_ERR_UNSUPPORTED when its rows would store more than
 * 1.5x nnz values; the Lanczos, Arnoldi, CG and multi-shift CG steps and
 * psfun_spmv read it) or PSFUN_FORMAT_AUTO (SELL when its padding stays within
 * 25% of nnz), or PSFUN_FORMAT_CSRS (CSR row blocks of at most 256 rows and
 * 1024 entries, a longer row alone: the Lanczos step streams each block's
 * entries with coalesced loads and sums its rows from LDS; the other kernels
 * read it as vector CSR).  Results differ only in rounding order.  (PSBLAS's
 * psb_d_hdia_sparse_mat is the reference ecosystem's counterpart.) */
int psfun_csr_set_format(psfun_csr_t A, int format);
/* Slice diagonals the diagonal-slice copy stores (values = 64 x that), 0 when
 * it has not been built. */
int psfun_csr_hdia_diagonals(psfun_csr_t A, int64_t *ndiag);
/* Format in use and the padded sliced-ELL entry count (0 when not built). */
int psfun_csr_get_format(psfun_csr_t A, int *format, int64_t *sell_entries);
/* Width of the column indices the Lanczos sliced-ELL step reads: 16 when
 * every column is within [-32768, 32767] of its row (16-bit offsets are kept
 * beside the 32-bit columns; results are the same), 32 otherwise, 0 without
 * a sliced-ELL copy. */
int psfun_csr_sell_index_bits(psfun_csr_t A, int *bits);

/* ---- device vectors ------------------------------------------------------ */
int psfun_vec_alloc(psfun_ctx_t ctx, int64_t n, double **dptr);
int psfun_vec_free(psfun_ctx_t ctx, double *dptr);
int psfun_vec_upload(psfun_ctx_t ctx, double *dptr, const double *host, int64_t n);
int psfun_vec_download(psfun_ctx_t ctx, double *host, const double *dptr, int64_t n);

/* ---- PSBLAS primitives --------------------------------------------------- */
/* y <- alpha*A*x + beta*y  (psb_spmm, psfun_d_lanczos.f90:198) */
int psfun_spmv(psfun_ctx_t ctx, psfun_csr_t A, double alpha, const double *x, double beta,
               double *y);
/* Y <- alpha*A*X + beta*Y for k column-major right-hand sides with leading
 * dimension ld (psb_spmm on a multivector). */
int psfun_spmm(psfun_ctx_t ctx, psfun_csr_t A, int k, double alpha, const double *X, int64_t ldx,
               double beta, double *Y, int64_t ldy);
/* Scalar results are returned on the host (psb_gedot / psb_genrm2 / psb_geamax). */
int psfun_dot(psfun_ctx_t ctx, int64_t n, const double *x, const double *y, double *result);
int psfun_nrm2(psfun_ctx_t ctx, int64_t n, const double *x, double *result);
int psfun_amax(psfun_ctx_t ctx, int64_t n, const double *x, double *result);
/* y <- alpha*x + beta*y  (psb_geaxpby) */
int psfun_axpby(psfun_ctx_t ctx, int64_t n, double alpha, const double *x, double beta, double *y);
/* x <- alpha*x  (v%scal) */
int psfun_scal(psfun_ctx_t ctx, int64_t n, double alpha, double *x);

/* ---- f(H) step on a small dense matrix (psfun_d_serial%apply) ------------ */
/* y = f(scaling*H) x for a k-by-k column-major host matrix H (k <= 512),
 * evaluated on the device (psfun_d_serial_apply_array,
 * serial/psfun_d_serial_mod.F90:172-523).  fname = PSFUN_FN_EXP: the variant
 * fixes how H is read, as the reference's EXPOKIT routines read it: CHBSYM as
 * symmetric from its upper triangle (DSCHBV), CHBHES as upper Hessenberg
 * (DNCHBV), CHBGEN/GENPADE/TAYLOR/SASQ as a full matrix.  fname =
 * PSFUN_FN_USERF, variant PSFUN_VAR_SYM (lines 420-485): the symmetric matrix
 * of H's upper triangle is diagonalised on the device (DSYEV('V','U') there),
 * f(scaling*lambda_i) is evaluated by the psfun_set_userf function, and
 * y = Q diag(f) Q^T x. */
int psfun_serial_apply_host(psfun_ctx_t ctx, int fname, int variant, double scaling, int k,
                            const double *H, const double *x, double *y);

/* ---- Krylov f(A)x (psfun_d_krylov%apply) --------------------------------- */
/* y = ||x|| V_k f(fscale*T_k) e1 by Lanczos with the reference's stopping rule,
 * one-step-lagged a-posteriori estimate and assembly (psfun_d_lanczos.f90:183-377).
 * The Krylov operator is opscale*A (the reference uses opscale = 1, ignoring
 * fun%scaling, psfun_d_lanczos.f90:172-173,198,239).  itmax >= 2, istop in {1,2}.
 * Outputs: *iter, *err; res (host, length >= itmax+1) receives res(1:itx+1)
 * and *nres = itx+1.  x and y are device vectors of length n.
 * fname = PSFUN_FN_USERF evaluates f(T) as the USERF/SYM variant does
 * (psfun_set_userf); with istop = 1 the loop then waits for the host once
 * per iteration, where the user function runs. */
int psfun_lanczos_apply(psfun_ctx_t ctx, psfun_csr_t A, int fname, double fscale, double opscale,
                        const double *x, double *y, double eps, int itmax, int istop, int *iter,
                        double *err, double *res, int *nres);

/* The same Lanczos apply without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * kernels are psfun_lanczos_apply's; the control block, partial sums and
 * Krylov basis live in a fixed 64 MiB device array of the library (one copy
 * per device, held until the process ends) instead of the context's
 * workspace.  Limit: 64 KiB + align256(control block) + 8 * align64(n) *
 * itmax bytes <= 64 MiB (n * itmax up to about 8.3 M, e.g. n = 160 000 at
 * itmax 50); larger inputs return PSFUN_ERR_TOO_LARGE before any work.
 * fname must be PSFUN_FN_EXP (USERF returns PSFUN_ERR_UNSUPPORTED).  Every
 * call on a device shares that device's array, so calls that overlap on one
 * device (two contexts, two host threads) race: serialise them.  Calls after
 * one another are fine (each clears what it reads first).  Results are
 * bitwise those of psfun_lanczos_apply; per-kernel timing is not recorded. */
int psfun_lanczos_apply_noalloc(psfun_ctx_t ctx, psfun_csr_t A, int fname, double fscale, double opscale,
                                const double *x, double *y, double eps, int itmax, int istop, int *iter,
                                double *err, double *res, int *nres);

/* y = ||x|| V_k f(fscale*H_k) e1 by Arnoldi (modified Gram-Schmidt) with the
 * reference's stopping rule and assembly (psfun_d_arnoldi.f90:183-329): the
 * Krylov operator is scaling*A (the reference multiplies its SpMV by
 * fun%scaling and evaluates f with scaling 1, lines 172-173, 202), f(H) is
 * read as `variant` reads a matrix (see psfun_serial_apply_host).  No
 * breakdown test, as in the reference; x = 0 gives y = 0 and iter = 0.
 * itmax in 2..512, istop in {1,2}; outputs as for psfun_lanczos_apply.
 * fname = PSFUN_FN_USERF with variant PSFUN_VAR_SYM: f(H_k) as the USERF/SYM
 * f(H) step reads H_k (the symmetric matrix of its upper triangle), with one
 * host wait per iteration when istop = 1. */
int psfun_arnoldi_apply(psfun_ctx_t ctx, psfun_csr_t A, int fname, int variant, double fscale, double scaling,
                        const double *x, double *y, double eps, int itmax, int istop, int *iter, double *err,
                        double *res, int *nres);

/* Shift-and-invert Arnoldi (krylov/rational/psfun_d_sai.f90:37-419): the
 * Krylov space of (A + tau I)^{-1}, each inner solve (A + tau I) w = v by
 * psfun_cg_shifted (prec NONE/DIAG, or BJAC built once per call on the
 * unshifted A; tolerance eps, inner_itmax, inner_istop;
 * A + tau I must be definite), modified Gram-Schmidt, and
 * y = ||x|| V_k f(fscale (H_k^{-1} - tau I)) e1 with the reference's
 * a-posteriori estimate from that same projection.  Differences from the
 * reference, which the reference's own comments describe as the intent:
 * the shift is applied to every diagonal entry (its loop `do j=1,i-1,i`
 * reaches only (1,1)), the inner solve starts from zero, and a run that does
 * not converge also assembles from f(H^{-1} - tau I) (the reference uses
 * f(H) there).  For tau = 0 and a converged run the results coincide.  f is
 * read as `variant` reads a matrix; GENPADE/CHBGEN/TAYLOR/SASQ treat the full
 * H^{-1}; fname = PSFUN_FN_USERF with PSFUN_VAR_SYM takes the symmetric
 * matrix of its upper triangle.  Returns PSFUN_ERR_BREAKDOWN if H_k is
 * singular. */
int psfun_sai_arnoldi_apply(psfun_ctx_t ctx, psfun_csr_t A, int fname, int variant, double fscale, double tau,
                            int prec, int inner_itmax, int inner_istop, const double *x, double *y, double eps,
                            int itmax, int istop, int *iter, double *err, double *res, int *nres);

/* ---- batched Lanczos f(A)x ------------------------------------------------ */
/* A batch is a set of ngraph independent square CSR matrices stored
 * block-diagonally: graph g owns rows row_offsets[g]..row_offsets[g+1]-1 of one
 * 0-based CSR (rowptr over all rows, colind GLOBAL and inside the graph's own
 * row range).  Host arrays; copied to the device.  No reference counterpart:
 * it runs ngraph*K independent psfun_d_lanczos instances
 * (krylov/polynomial/psfun_d_lanczos.f90:37-400) in one launch sequence. */
typedef struct psfun_batch *psfun_batch_t;
int psfun_batch_create(psfun_ctx_t ctx, int ngraph, const int64_t *row_offsets, const int32_t *rowptr,
                       const int32_t *colind, const double *val, psfun_batch_t *B);
/* Row order a batch keeps on the device.  RCM renumbers each graph's rows by
 * reverse Cuthill-McKee (psfun_rcm_order) so that the neighbour gathers of
 * the SpMM stay close together; X, Y and values stay in the caller's
 * numbering (the batch gathers X, scatters Y and permutes values itself), and
 * each row keeps its entries in the given order, so results equal the
 * natural order's up to the order of the dot-product reductions.  Calls that
 * would expose the internal numbering (psfun_batch_csr,
 * psfun_batch_lanczos_basis) return PSFUN_ERR_UNSUPPORTED on such a batch. */
enum { PSFUN_ORDER_NATURAL = 0, PSFUN_ORDER_RCM = 1 };
int psfun_batch_create_ordered(psfun_ctx_t ctx, int ngraph, const int64_t *row_offsets, const int32_t *rowptr,
                               const int32_t *colind, const double *val, int order, psfun_batch_t *B);
/* Reverse Cuthill-McKee order of one graph (host): n rows, rowptr[0..n] into
 * colind, column indices base..base+n-1; perm[p] = original local row of new
 * row p.  No device work. */
int psfun_rcm_order(int n, const int32_t *rowptr, const int32_t *colind, int base, int32_t *perm);
int psfun_batch_destroy(psfun_batch_t B);
/* Tuning knobs (set by the autotuner): rows per block-chunk, and whether the
 * whole launch sequence is captured once in a hipGraph and replayed. */
int psfun_batch_set_tuning(psfun_batch_t B, int rows_per_chunk, int use_graph);
int psfun_batch_info(psfun_batch_t B, int *ngraph, int64_t *nrows, int64_t *nnz, int *rows_per_chunk, int *nchunks);
/* Y = exp(fscale * opscale*A_g) X for every graph g and each of the K columns of
 * the row-major (nrows x K, K <= 64) device blocks X, Y, with the reference's
 * Lanczos stopping rule per instance.  iters/errs (host, ngraph*K, instance
 * g*K+k; may be NULL) receive iter and err of every instance. */
int psfun_batch_lanczos(psfun_ctx_t ctx, psfun_batch_t B, int K, double fscale, double opscale, const double *X,
                        double *Y, double eps, int itmax, int istop, int *iters, double *errs);

/* Y = alpha*A_blockdiag*X + beta*Y on row-major (nrows x K) device blocks. */
int psfun_batch_spmm(psfun_ctx_t ctx, psfun_batch_t B, int K, double alpha, const double *X, double beta, double *Y);
/* Replace the nonzero values (same pattern) from a device array of nnz doubles. */
int psfun_batch_set_values(psfun_batch_t B, const double *val);
/* Device pointers of the batch's 0-based block-diagonal CSR arrays. */
int psfun_batch_csr(psfun_batch_t B, const int32_t **rowptr, const int32_t **colind, const double **val);

/* As psfun_batch_lanczos, but the Krylov basis goes to the caller's device
 * buffer V ((itmax-1)*nrows*K doubles; basis slot l of every instance at
 * V + l*nrows*K, row-major like X) and each instance's Lanczos data to caller
 * device arrays: T_alpha/T_beta (ninst*(itmax+2); instance I at I*(itmax+2),
 * 1-based as T(i,i) and T(i,i+1)), nbasis (ninst ints: valid basis vectors)
 * and rs1 (ninst: ||x||).  The basis is kept unnormalised: slot 0 holds x and
 * slot l >= 1 holds w_l = T(l,l+1) v_{l+1}, so v_{l+1} = slot l / T_beta(l)
 * (v_1 = x / rs1).  Used to reuse the basis in an adjoint pass. */
int psfun_batch_lanczos_basis(psfun_ctx_t ctx, psfun_batch_t B, int K, double fscale, double opscale, const double *X,
                              double *Y, double eps, int itmax, int istop, double *V, double *T_alpha, double *T_beta,
                              int *nbasis, double *rs1);

/* ---- adjoint helpers (differentiable exp(-tA)X; no reference counterpart) - */
/* Galerkin resolvents from a stored Krylov basis (one matrix, K columns):
 * out_s[r, col] = ||x_col|| (V_m (xi_s I + T_m)^{-1} e1)[r] for nshift <= 32
 * host complex shifts xi_s; out is device complex, entry (s, r, col) at
 * out + 2*((s*nrows + r)*K + col).  V, T_alpha, T_beta, nbasis, rs1 as written
 * by psfun_batch_lanczos_basis (ldt = itmax+2, mmax >= max nbasis); the stored
 * tridiagonal belongs to the Lanczos operator opscale*A and tscale = 1/opscale
 * maps it to A. */
int psfun_krylov_resolvents(psfun_ctx_t ctx, int64_t nrows, int K, int mmax, const double *V, const double *T_alpha,
                            const double *T_beta, int ldt, double tscale, const int *nbasis, const double *rs1,
                            int nshift, const double *shifts, double *out);
/* grad[q] += scale * sum_s Re(w_s sum_col U_s[i,col] V_s[k,col]) for every
 * nonzero q = (i, k) of the device CSR pattern; U, V laid out as the output
 * of psfun_krylov_resolvents; host weights w (nshift <= 32). */
int psfun_sddmm_cols(psfun_ctx_t ctx, int n, const int32_t *rowptr, const int32_t *colind, int K, int nshift,
                     const double *weights, const double *U, const double *V, double scale, double *grad);
/* One-call vector-Jacobian product of Y = exp(-tA) X for a symmetric A held
 * as the single graph of a batch in natural order (no reference counterpart:
 * the differentiable form of the f(A)x path for C and Fortran callers).  For
 * dL/dY = G (device, nrows x K row-major like X and Y, K <= 64):
 *   gX    (device, nrows x K)  = exp(-tA) G                  (may be NULL)
 *   gt    (host scalar)        = -<G, A Y>; needs Y          (may be NULL)
 *   gvals (device, nnz)        = dL/d(stored nonzeros of A), overwritten; needs
 *          X and t > 0; from the npoles-point modified Talbot resolvent form
 *          with conjugate poles folded, resolvents taken from the Lanczos
 *          bases of G and of X (the forward basis is recomputed) (may be NULL)
 * Lanczos runs use eps, itmax and istop 1.  Allocates its bases and the
 * resolvents ((itmax-1)*nrows*K + 4*S*nrows*K doubles, S <= 32 folded poles)
 * and returns after the stream has finished. */
int psfun_expv_vjp(psfun_ctx_t ctx, psfun_batch_t B, int K, double t, const double *X, const double *Y,
                   const double *G, double eps, int itmax, int npoles, double *gX, double *gt, double *gvals);

/* ---- shifted solves (psb_krylov extension) -------------------------------- */
/* Preconditioners of the shifted CG (psb_dprec_type%init(ctxt, ptype, info),
 * applied at krylov/shifted/psfun_dcg.F90:283): NONE, DIAG (inverse diagonal
 * of the unshifted A) and BJAC (block Jacobi, which with one process is the
 * whole matrix factored by ILU(0): M = L U on A's sparsity pattern). */
enum { PSFUN_PREC_NONE = 0, PSFUN_PREC_DIAG = 1, PSFUN_PREC_BJAC = 2 };
typedef struct psfun_prec *psfun_prec_t;
/* Build a preconditioner on the (unshifted) matrix A, as prec%build(a,
 * desc_a, info) does: BJAC factors A by ILU(0) on the host (rows sorted by
 * column, repeated entries summed; a zero pivot returns PSFUN_ERR_BREAKDOWN)
 * and uploads the factors with the level schedules of their triangular
 * solves; NONE and DIAG keep no state.  The object may outlive its context
 * only to be destroyed (as matrices). */
int psfun_prec_create(psfun_ctx_t ctx, psfun_csr_t A, int kind, psfun_prec_t *P);
int psfun_prec_destroy(psfun_prec_t P);
/* kind and the number of level-scheduled steps of one BJAC application */
int psfun_prec_info(psfun_prec_t P, int *kind, int *levels);
/* z = M^{-1} r (BJAC: U^{-1} L^{-1} r), device vectors; PSFUN_ERR_UNSUPPORTED
 * for NONE/DIAG (the solver applies those itself). */
int psfun_prec_apply(psfun_ctx_t ctx, psfun_prec_t P, const double *r, double *z);
/* (eta*A + zeta*I) x = b by preconditioned CG with the semantics of
 * psfun_dcg_vect (krylov/shifted/psfun_dcg.F90:129-373): x holds the initial
 * guess on entry; prec NONE, DIAG (diagonal of the unshifted A) or BJAC
 * (ILU(0) of the unshifted A, built for the call); istop 1
 * (||r||_inf <= eps(||A||_inf ||x||_inf + ||b||_inf)) or 2 (||r||_2 <= eps ||b||_2);
 * breakdowns restart as in the reference.  cond (may be NULL) receives the
 * extreme-eigenvalue ratio of the CG tridiagonal (lines 314-354). */
int psfun_cg_shifted(psfun_ctx_t ctx, psfun_csr_t A, int prec, const double *b, double eta, double zeta,
                     double *x, double eps, int itmax, int istop, int *iter, double *err, double *cond);
/* The same solver with a built preconditioner object (the reference's
 * prec%apply): prec = PSFUN_PREC_BJAC builds one per call in
 * psfun_cg_shifted, this entry reuses P across solves. */
int psfun_cg_shifted_prec(psfun_ctx_t ctx, psfun_csr_t A, psfun_prec_t P, const double *b, double eta,
                          double zeta, double *x, double eps, int itmax, int istop, int *iter, double *err,
                          double *cond);
/* Multi-shift CG: (A + sigma_s I) X(:,s) = b for nshift <= 64 complex shifts
 * (host array of (re,im) pairs), A symmetric positive definite, one shared
 * real Krylov space.  X is device complex (interleaved), column s at
 * X + 2*s*ldx.  Shift s stops when its residual is <= eps*||b||; iters/resid
 * (host, may be NULL) receive each shift's iteration count and final relative
 * residual.  No reference counterpart (psfun_dcg.F90 solves one real shift). */
int psfun_mscg(psfun_ctx_t ctx, psfun_csr_t A, int nshift, const double *shifts, const double *b, double *X,
               int64_t ldx, double eps, int itmax, int *iters, double *resid);

/* ---- quadrature rules and special functions (host, no device needed) ------ */
/* f evaluated at the complex point z[0] + i z[1]; result to fz[0..1]. */
typedef void (*psfun_zfun_t)(const double *z, double *fz, void *user);
/* sum_{i<nc} c[i] x^i  (horner, utils/psfun_d_utils_mod.f90:119-138) */
double psfun_horner(int nc, const double *c, double x);
/* K(m), K(1-m) at m = exp(-2 pi L)  (ellipkkp, utils/psfun_d_utils_mod.f90:37-99) */
int psfun_ellipkkp(double L, double *K);
/* Jacobi sn, cn, dn at complex u, m = exp(-2 pi L); out = (sn, cn, dn) as
 * (re, im) pairs  (ellipj, utils/psfun_z_utils_mod.f90:39-107) */
int psfun_ellipj(double u_re, double u_im, double L, double *out);
/* Hale-Higham-Trefethen method 1 poles xi(N), weights c(N) (complex pairs),
 * eta, sign for f on the spectrum interval rparams[0..1]
 * (hhtmethod1, quadratures/psfun_z_quadrules_mod.f90:40-85). */
int psfun_hhtmethod1(psfun_zfun_t zfun, void *user, int N, const double *rparams, double *xi, double *c,
                     double *eta, double *sign);
/* Modified Talbot rule for exp(-t z): exp(-tA)x ~ Re sum_j c_j (xi_j I + A)^{-1} x
 * (eta = 1, sign = +1 in the psfun_z_quadrature convention).  No reference
 * counterpart: the reference ships only hhtmethod1. */
int psfun_talbot(int N, double t, double *xi, double *c, double *eta, double *sign);

/* ---- rational quadrature apply (psfun_z_quadrature) ----------------------- */
/* v_s = (A + sigma_s I)^{-1} x by psfun_mscg, then
 *   form PSFUN_QUAD_RE_SUM:   y = gamma * Re(sum_s w_s v_s)
 *   form PSFUN_QUAD_A_IM_SUM: y = gamma * A * Im(sum_s w_s v_s)   (HHT method 1)
 * The reference stores poles/weights only (quadratures/psfun_z_quadrature_mod.F90:
 * 42-55, 114-142); this is the apply its documentation describes
 * (docs/pages/algorithms.md:106-125). */
enum { PSFUN_QUAD_RE_SUM = 0, PSFUN_QUAD_A_IM_SUM = 1 };
/* Preconditioned forms of psfun_mscg and psfun_quad_apply: the setting
 * psfun_z_quadrature%set(prec) stores for the poles' solves
 * (quadratures/psfun_z_quadrature_mod.F90:156-166), applied as
 * krylov/shifted/psfun_dcg.F90:283 applies prec%apply.  A preconditioner of A
 * is not shift-invariant, so the multi-shift solve runs first and every shift
 * it leaves above eps after itmax iterations continues from its iterate by
 * preconditioned COCG (all such shifts in one launch sequence) for at most
 * itmax more iterations, with M = I (P of kind NONE), diag(A) + sigma_s I
 * (DIAG) or the ILU(0) of A (BJAC, applied to the real and imaginary parts).
 * P = NULL is psfun_mscg / psfun_quad_apply.  iters = multi-shift plus COCG
 * iterations; resid = the final ||r_s|| / ||b|| (recursive estimate for the
 * shifts the multi-shift solve finished, true residual for the others).
 * nshift <= 64.  The quadrature form keeps the shifted solutions in the
 * context's outer workspace (16 n nshift bytes). */
int psfun_mscg_prec(psfun_ctx_t ctx, psfun_csr_t A, psfun_prec_t P, int nshift, const double *shifts,
                    const double *b, double *X, int64_t ldx, double eps, int itmax, int *iters, double *resid);
int psfun_quad_apply_prec(psfun_ctx_t ctx, psfun_csr_t A, psfun_prec_t P, int nshift, const double *shifts,
                          const double *weights, double gamma, int form, const double *x, double *y, double eps,
                          int itmax, int *iters, double *resid);
/* Sampled derivative for the adjoint of a resolvent sum: for every nonzero q
 * = (i, k) of the device CSR pattern (rowptr, colind; n rows),
 *   grad[q] += scale * sum_s Re(w_s U_s[i] V_s[k]),
 * with U_s, V_s device complex vectors (interleaved) at U + 2*s*ldu, V + 2*s*ldv
 * and host weights w (nshift <= 64).  No reference counterpart. */
int psfun_sddmm_shifted(psfun_ctx_t ctx, int n, const int32_t *rowptr, const int32_t *colind, int nshift,
                        const double *weights, const double *U, int64_t ldu, const double *V, int64_t ldv,
                        double scale, double *grad);
int psfun_quad_apply(psfun_ctx_t ctx, psfun_csr_t A, int nshift, const double *shifts, const double *weights,
                     double gamma, int form, const double *x, double *y, double eps, int itmax, int *iters,
                     double *resid);

#ifdef __cplusplus
}
#endif

#endif /* PSFUN_HIP_H */
