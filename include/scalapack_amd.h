/*
 * scalapack_amd.h -- C-ABI of the MI355X (gfx950) dense LU solve library.
 *
 * Part 1 is the drop-in: the Fortran-callable ScaLAPACK entry points of the
 * LU solve path, with the reference's exact calling sequence (every argument
 * by reference, 9-integer descriptors, INFO semantics, gfortran/flang
 * "Add_" name mangling as in the reference's SLmake.inc.example:21 and
 * CMAKE/FortranMangling.cmake:42).  A Fortran program that calls PDGESV /
 * PDGETRF / PDGETRS links against libscalapack_amd.so instead of (or ahead
 * of) libscalapack and runs them on the GPU for a 1 x 1 BLACS grid.  Arrays
 * are host memory, as in the reference; they are staged to HBM inside the
 * call.  Character arguments carry the hidden length Fortran compilers
 * append; it is accepted and ignored, as the reference's C PBLAS does.
 *
 * Part 2 is a device-resident interface (pointers into HBM) for callers that
 * keep the matrix on the GPU; the benchmark and the GPU tests use it.
 *
 * The 1 x 1 BLACS context routines the drop-in needs (BLACS_GRIDINFO etc.)
 * come from libsla_blacs1x1.so (sla_blacs1x1.h) unless a real BLACS is linked
 * ahead of it.
 */
#ifndef SCALAPACK_AMD_H
#define SCALAPACK_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Part 1: drop-in ScaLAPACK entry points (Fortran ABI, 32-bit INTEGER).     */
/* ------------------------------------------------------------------------ */

/* PDGESV: solve sub(A) * X = sub(B).  Replaces SRC/pdgesv.f:1-2
 * (checks :185-225, PDGETRF :229, PDGETRS :236). */
void pdgesv_(const int *n, const int *nrhs, double *a, const int *ia, const int *ja,
             const int *desca, int *ipiv, double *b, const int *ib, const int *jb,
             const int *descb, int *info);

/* PDGETRF: LU factorization with partial pivoting, sub(A) = P * L * U.
 * Replaces SRC/pdgetrf.f:1 (checks :169-199, quick return :201-206,
 * blocked loop :217-295).  IPIV(i) = global row interchanged with row i. */
void pdgetrf_(const int *m, const int *n, double *a, const int *ia, const int *ja,
              const int *desca, int *ipiv, int *info);

/* PDGETRS: solve with the factors of PDGETRF.  Replaces SRC/pdgetrs.f:1-2
 * (checks :184-237; TRANS = 'N': PDLAPIV :255, PDTRSM :260,265; TRANS = 'T'
 * or 'C': PDTRSM :276,279 with U^T and L^T, PDLAPIV backward :282). */
void pdgetrs_(const char *trans, const int *n, const int *nrhs, const double *a, const int *ia,
              const int *ja, const int *desca, const int *ipiv, double *b, const int *ib,
              const int *jb, const int *descb, int *info, size_t trans_len);

/* PDGETF2: unblocked panel factorization, N + MOD(JA-1, NB) <= NB.
 * Replaces SRC/pdgetf2.f:1 (checks :170-193: an illegal argument prints the
 * PXERBLA message and aborts, as BLACS_ABORT does there). */
void pdgetf2_(const int *m, const int *n, double *a, const int *ia, const int *ja,
              const int *desca, int *ipiv, int *info);

/* PDLASWP: row ('R') or column ('C') interchanges K1..K2, forward ('F') or
 * backward ('B').  Replaces SRC/pdlaswp.f:1-2 (:159-203). */
void pdlaswp_(const char *direc, const char *rowcol, const int *n, double *a, const int *ia,
              const int *ja, const int *desca, const int *k1, const int *k2, const int *ipiv,
              size_t direc_len, size_t rowcol_len);

/* PDLANGE: 'M' max |a_ij|, '1'/'O' one norm, 'I' infinity norm, 'F'/'E'
 * Frobenius norm of sub(A).  Replaces SRC/pdlange.f:1-2 (:196-330); WORK is
 * not referenced.  The one and infinity norms sum in the reference's order. */
double pdlange_(const char *norm, const int *m, const int *n, const double *a, const int *ia,
                const int *ja, const int *desca, double *work, size_t norm_len);

/* PDGECON: reciprocal condition number estimate from the factors of
 * PDGETRF.  Replaces SRC/pdgecon.f:1-3 (checks :215-285, quick returns
 * :287-297, PDLACON loop :324-389 with PDLATRS = triangular solves).  LWORK /
 * LIWORK minimums and the LWORK = -1 query as the reference. */
void pdgecon_(const char *norm, const int *n, const double *a, const int *ia, const int *ja,
              const int *desca, const double *anorm, double *rcond, double *work, const int *lwork,
              int *iwork, const int *liwork, int *info, size_t norm_len);

/* PDGERFS: iterative refinement with forward / backward error bounds.
 * Replaces SRC/pdgerfs.f:1-4 (checks :308-404, quick return :440-446,
 * refinement :480-556, error bound :557-666). */
void pdgerfs_(const char *trans, const int *n, const int *nrhs, const double *a, const int *ia,
              const int *ja, const int *desca, const double *af, const int *iaf, const int *jaf,
              const int *descaf, const int *ipiv, const double *b, const int *ib, const int *jb,
              const int *descb, double *x, const int *ix, const int *jx, const int *descx,
              double *ferr, double *berr, double *work, const int *lwork, int *iwork,
              const int *liwork, int *info, size_t trans_len);

/* PDGEEQU: row and column scalings that equilibrate sub(A).  Replaces
 * SRC/pdgeequ.f:1-2; R(IA..IA+M-1), C(JA..JA+N-1); INFO = i (row i of sub(A)
 * exactly zero) or M + j (column j). */
void pdgeequ_(const int *m, const int *n, const double *a, const int *ia, const int *ja,
              const int *desca, double *r, double *c, double *rowcnd, double *colcnd, double *amax,
              int *info);

/* PDLAQGE: applies the scalings of PDGEEQU when they are needed (THRESH =
 * 0.1).  Replaces SRC/pdlaqge.f:1-2; EQUED = 'N', 'R', 'C' or 'B'. */
void pdlaqge_(const int *m, const int *n, double *a, const int *ia, const int *ja, const int *desca,
              const double *r, const double *c, const double *rowcnd, const double *colcnd,
              const double *amax, char *equed, size_t equed_len);

/* PDGESVX: the expert driver -- equilibration (FACT = 'E'), PDGETRF (unless
 * FACT = 'F'), PDGECON, PDGETRS, PDGERFS and the scaling of X and FERR back.
 * Replaces SRC/pdgesvx.f:1-5 (checks :452-620, steps :624-780); INFO = IA+N
 * when RCOND < eps (X not computed). */
void pdgesvx_(const char *fact, const char *trans, const int *n, const int *nrhs, double *a,
              const int *ia, const int *ja, const int *desca, double *af, const int *iaf,
              const int *jaf, const int *descaf, int *ipiv, char *equed, double *r, double *c,
              double *b, const int *ib, const int *jb, const int *descb, double *x, const int *ix,
              const int *jx, const int *descx, double *rcond, double *ferr, double *berr,
              double *work, const int *lwork, int *iwork, const int *liwork, int *info,
              size_t fact_len, size_t trans_len, size_t equed_len);

/* ------------------------------------------------------------------------ */
/* Part 2: device-resident interface.                                        */
/* ------------------------------------------------------------------------ */

typedef struct sla_handle_s *sla_handle_t;

/* Return codes of the sla_* functions: 0 success, SLA_ERR_* on failure
 * (the message is in sla_last_error()).  INFO-style results (a zero pivot)
 * are returned through an `info` argument, never as an error. */
#define SLA_OK 0
#define SLA_ERR_ARG (-1)
#define SLA_ERR_RUNTIME (-2)

/* Creates a handle bound to the current HIP device. */
int sla_create(sla_handle_t *h);
int sla_destroy(sla_handle_t h);
/* Work of the handle goes to `stream` (a hipStream_t; NULL is HIP's default
 * stream, as for rocBLAS).  A new handle uses a non-blocking stream of its
 * own: inputs produced on other streams must be complete before a call.
 * Calls return after the stream has drained. */
int sla_set_stream(sla_handle_t h, void *stream);
/* Tuning knobs: "nb" (block column width, 1..512), "leaf_w" (panel leaf width,
 * <= 32), "panel_wgs" (workgroups per panel leaf, 1..128), "profile" (0/1); for the
 * mixed-precision solve "ir_max_iter" (3..1000, default 30), "ir_gmres_on_stall"
 * (0/1, default 1), "ir_accept_stall" (0/1, default 1: a refinement that
 * stalls with scaled residual < 16 returns its X; 0: only DSGESV's
 * convergence test ends refinement, a stall falls back to FP64), "ir_fallback"
 * (0/1, default 1: a failed refinement is redone by the FP64 factorization;
 * 0: the failure is reported with ITER < 0 and SLA_STOP_FAILED, A untouched --
 * for callers timing the low-precision attempt alone), "trsv_blockinv" (0/1,
 * default 1: solves with low-precision factors multiply by the inverses of
 * their 128 x 128 diagonal blocks, formed once per factorization; 0:
 * substitution), "trsm_mfma" (0/1, default 1: the block row U12 <- L11^{-1} A12
 * of every block column is one launch with the 32 x 32 diagonal blocks
 * applied by their inverses and the products on MFMA; 0: recursive 32-row
 * substitution strips and GEMMs), "gmres_m" (GMRES steps per correction, default 50),
 * "gmres_tol_exp" (e: relative reduction 10^-e per correction, default 6),
 * "pivoting" (FP32-storage factorizations: 0 partial pivoting, 1 tournament
 * pivoting of the panel leaves, default 1; the FP64 path always pivots as
 * PDGETF2), "lookahead" (0/1, default 1: block column p+1 is factored on a
 * second stream during p's update, every row interchange stays on the
 * caller's stream; 0: one stream -- the same arithmetic in serial launch
 * order, bit-identical factors and IPIV), "update_cus" (0 = all; else the trailing
 * update runs as a persistent grid sized for this many CUs, leaving the
 * others to the panel), "lacon_lapack" (0/1, default 0: the condition and
 * forward-error estimates of sla_dgecon / sla_dgerfs / pdgecon_ / pdgerfs_
 * return what the reference's PDLACON returns; 1: LAPACK DLACON's
 * iteration, never smaller and usually closer to the true norm),
 * "update_lds_kb" (0..160, default 112: LDS reserved per workgroup of the FP32
 * trailing update so that one of its workgroups occupies a CU and the panel
 * leaf running beside it finds registers; 0 = none). */
int sla_set_param(sla_handle_t h, const char *name, int value);
int sla_get_param(sla_handle_t h, const char *name, int *value);
const char *sla_last_error(void);

/* Per-kernel-class timing collected while "profile" is 1: classes
 * 0 panel, 1 laswp, 2 trsm, 3 gemm (the outer trailing updates), 4 trsv,
 * 5 other (conversions, residuals, packing), 6 the GEMMs inside the recursive
 * panel and TRSM.  ms = summed kernel time,
 * work = summed algorithmic flops (bytes for laswp), launches = count. */
int sla_get_stats(sla_handle_t h, int cls, double *ms, double *work, long long *launches);
/* The outer trailing-update launches of class 3 as the kernels themselves
 * measure them while "profile" is 1 (no reference counterpart: a measurement
 * aid): ms = summed execution intervals from the first workgroup's start to
 * the last workgroup's end on the device wall clock (the stream's event pair
 * also counts the wait for CUs held by the look-ahead panel stream),
 * work = summed 2 m n k flops of the 128 x 128-tiled interiors. */
int sla_get_update_exec(sla_handle_t h, double *ms, double *work, long long *launches);
int sla_reset_stats(sla_handle_t h);

/* FP64 LU of the m x n column-major matrix dA (leading dimension lda) in
 * HBM.  dipiv: device int[min(m,n)], 1-based pivot rows (PDGETRF's IPIV for
 * IA = JA = 1).  *info as PDGETRF's INFO (> 0: exactly zero pivot). */
int sla_dgetrf(sla_handle_t h, int m, int n, double *dA, int lda, int *dipiv, int *info);
/* Solve A X = B with the factors of sla_dgetrf (TRANS = 'N'). */
int sla_dgetrs(sla_handle_t h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
               double *dB, int ldb);
int sla_dgesv(sla_handle_t h, int n, int nrhs, double *dA, int lda, int *dipiv, double *dB,
              int ldb, int *info);
/* Solve op(A) X = B, op(A) = A ('N') or A^T ('T', 'C'), with the factors of
 * sla_dgetrf (PDGETRS with TRANS, SRC/pdgetrs.f:255-284). */
int sla_dgetrs_trans(sla_handle_t h, char trans, int n, int nrhs, const double *dA, int lda,
                     const int *dipiv, double *dB, int ldb);
/* PDGEEQU / PDLAQGE / PDGESVX on device data (dR, dC: device doubles). */
int sla_dgeequ(sla_handle_t h, int m, int n, const double *dA, int lda, double *dR, double *dC,
               double *rowcnd, double *colcnd, double *amax, int *info);
int sla_dlaqge(sla_handle_t h, int m, int n, double *dA, int lda, const double *dR,
               const double *dC, double rowcnd, double colcnd, double amax, char *equed);
/* fact 'N', 'E' or 'F'; *equed in ('F') / out; ferr, berr host arrays;
 * *info = n + 1 when RCOND < eps. */
int sla_dgesvx(sla_handle_t h, char fact, char trans, int n, int nrhs, double *dA, int lda,
               double *dAF, int ldaf, int *dipiv, char *equed, double *dR, double *dC, double *dB,
               int ldb, double *dX, int ldx, double *rcond, double *ferr, double *berr, int *info);
/* PDLANGE of an m x n device matrix (SRC/pdlange.f:196-330); norm one of
 * 'M', '1', 'O', 'I', 'F', 'E'. */
int sla_dlange(sla_handle_t h, char norm, int m, int n, const double *dA, int lda,
               double *value);
/* PDGECON with the factors of sla_dgetrf (SRC/pdgecon.f:287-403): *rcond =
 * 1 / (anorm * estimate of ||A^{-1}||) in the '1' / 'O' or 'I' norm. */
int sla_dgecon(sla_handle_t h, char norm, int n, const double *dLU, int lda, double anorm,
               double *rcond);
/* PDGERFS (SRC/pdgerfs.f:473-675): refines the solutions dX of op(A) X = dB
 * given A (dA) and its factors (dAF, dipiv from sla_dgetrf); ferr / berr are
 * host arrays of nrhs doubles (ferr[k] is left as it was when x_k is zero). */
int sla_dgerfs(sla_handle_t h, char trans, int n, int nrhs, const double *dA, int lda,
               const double *dAF, int ldaf, const int *dipiv, const double *dB, int ldb, double *dX,
               int ldx, double *ferr, double *berr);

/* The update GEMM alone: C[m x n] -= A[m x k] * B[k x n] (FP64,
 * column-major, device pointers) -- PDGEMM's local dgemm_('N','N', ..., -1,
 * L21, U12, +1, A22) of SRC/pdgetrf.f:288 (PBLAS/SRC/PTOOLS/PB_CpgemmAB.c:
 * 333-352), through the same shape dispatch as the factorization's GEMMs.
 * For kernel-level checks and measurement. */
int sla_dgemm_sub(sla_handle_t h, int m, int n, int k, const double *dA, int lda,
                  const double *dB, int ldb, double *dC, int ldc);

/* ------------------------------------------------------------------------ */
/* Part 3: mixed precision (no counterpart in the reference; the contract is  */
/* that of LAPACK's DSGESV, the mixed-precision form of PDGESV's solve).      */
/* ------------------------------------------------------------------------ */

/* Factorization precisions: FP64 (plain), or FP32 storage whose trailing
 * updates run with FP32 accumulation on
 *   FP32 MFMA (SLA_PREC_FP32), BF16 operands (_BF16), FP8-E4M3 operands with
 *   per-block-column scales (_FP8), or split 16-bit products -- each FP32
 *   operand as a sum of 16-bit slices, the slice products of low order kept:
 *   _BF16X3 (2 BF16 slices, 3 products, ~2^-16 |a||b|), _BF16X6 (3 slices,
 *   6 products, ~2^-24: FP32's own level), _FP16X3 (2 FP16 slices with a
 *   power-of-two scale per operand block, 3 products, ~2^-21). */
#define SLA_PREC_FP64 0
#define SLA_PREC_FP32 1
#define SLA_PREC_BF16 2
#define SLA_PREC_FP8 3
#define SLA_PREC_BF16X3 4
#define SLA_PREC_BF16X6 5
#define SLA_PREC_FP16X3 6
/* Refinement: classical iterative refinement, or GMRES-IR (flexible GMRES on
 * A d = r preconditioned by the low-precision factors). */
#define SLA_REFINE_IR 0
#define SLA_REFINE_GMRES 1

typedef struct {
    int precision;   /* factorization precision asked for (SLA_PREC_*) */
    int refinement;  /* refinement that produced the result (SLA_REFINE_*) */
    int iter;        /* DSGESV's ITER: refinement steps; < 0 after FP64 fallback */
    int outer;       /* FP64 residuals computed after the first solve */
    int inner;       /* GMRES steps in total */
    int fell_back;   /* 1 if the FP64 factorization produced X */
    double anorm;    /* ||A||_inf */
    double bwd_err;  /* max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf); NaN after
                        the FP64 fallback (A then holds the factors) */
    double sresid;   /* bwd_err / (N eps), eps = 2^-53: the testers' scaled residual */
    double history[16]; /* bwd_err after the first solve and each refinement step */
    int stop_rule;   /* what ended the refinement: SLA_STOP_CONVERGED (DSGESV's test),
                        SLA_STOP_STALL_ACCEPTED (a stall, PDGERFS's rule, with scaled
                        residual < 16; only while "ir_accept_stall" is 1) or
                        SLA_STOP_FALLBACK (X from the FP64 factorization),
                        SLA_STOP_FAILED (refinement failed and the handle's
                        "ir_fallback" is 0: X is the last iterate, A untouched) */
} sla_solve_info_t;
#define SLA_STOP_CONVERGED 0
#define SLA_STOP_STALL_ACCEPTED 1
#define SLA_STOP_FALLBACK 2
#define SLA_STOP_FAILED 3

/* X = A^{-1} B (n x n FP64 dA, nrhs <= 16, device pointers) by a
 * factorization in precision `prec` and refinement `refine` to FP64 backward
 * error: ||r||_inf <= ||x||_inf ||A||_inf sqrt(n) eps per column, or a stall
 * below the scaled-residual threshold 16.  dA is unchanged unless the FP64
 * fallback ran (then it holds L\U and dipiv the pivots, as sla_dgetrf);
 * dB is not modified.  *iter as DSGESV's ITER (-1 prec was FP64, -2 overflow
 * in FP32, -3 zero pivot in low precision, -(ir_max_iter+1) no convergence);
 * *info as PDGESV's INFO. */
int sla_dxgesv(sla_handle_t h, int prec, int refine, int n, int nrhs, double *dA, int lda,
               int *dipiv, const double *dB, int ldb, double *dX, int ldx, int *iter, int *info);
/* Details of the last sla_dxgesv on this handle. */
int sla_last_solve_info(sla_handle_t h, sla_solve_info_t *out);

/* FP32-storage LU (the low-precision factorization alone), trailing updates
 * in `prec` (SLA_PREC_FP32 .. SLA_PREC_FP16X3).  Pivots and *info as sla_dgetrf. */
int sla_sgetrf(sla_handle_t h, int prec, int m, int n, float *dA, int lda, int *dipiv,
               int *info);
/* Solve A X = B with the FP32 factors of sla_sgetrf (FP32 right-hand sides). */
int sla_sgetrs(sla_handle_t h, int n, int nrhs, const float *dA, int lda, const int *dipiv,
               float *dB, int ldb);
/* The refinement's FP64 residual alone: R = B - A X (n x n A, device
 * pointers), rnorm[j] = ||R(:, j)||_inf (host array of nrhs doubles).  The
 * sum order is fixed, so R is reproducible bit for bit. */
int sla_dresidual(sla_handle_t h, int n, int nrhs, const double *dA, int lda, const double *dX,
                  int ldx, const double *dB, int ldb, double *dR, int ldr, double *rnorm);
/* ------------------------------------------------------------------------ */
/* Part 4: batched small systems (no counterpart in the reference: thousands  */
/* of independent PDGESV-sized-down solves in persistent launches).          */
/* ------------------------------------------------------------------------ */
/* A_s X_s = B_s for s < batch, each as PDGESV on a 1 x 1 grid (partial
 * pivoting with IDAMAX's first maximum, rows interchanged across the matrix,
 * reciprocal scaling: SRC/pdgetf2.f:206-237; solve as SRC/pdgetrs.f:255-266,
 * skipped when INFO > 0 as SRC/pdgesv.f:231).  prec: SLA_PREC_FP64 (double
 * matrices) or SLA_PREC_FP32 (float).  n, nrhs, lda, ldb: host int[batch]
 * (n <= 1024, nrhs <= 16); dA, dB: device arrays of device pointers to the
 * column-major matrices; dipiv: device array of device pointers to int[n]
 * (1-based rows); dinfo: device int[batch].  One persistent launch factors the
 * systems largest first, a second one solves them (a kernel of its own, so
 * its registers do not lower the factorization's occupancy).  Handle knobs: "batched_w" (panel width 8 / 16 /
 * 32), "batched_wgs_per_cu" (1..8), "batched_graph" (1: the launch is
 * captured into a hipGraph once per set of pointer arrays and sizes, then
 * replayed). */
int sla_gesv_batched(sla_handle_t h, int prec, int batch, const int *n, const int *nrhs,
                     void *const *dA, const int *lda, int *const *dipiv, void *const *dB,
                     const int *ldb, int *dinfo);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sla_gesv_batched with the sizes in DEVICE arrays dn,
 * dnrhs, dlda, dldb and no allocation of any kind -- the queue, the order and
 * the size descriptors live in a fixed device array of the library (about
 * 1.3 MB per GPU, held until the process ends), so at most
 * SLA_BATCHED_NOALLOC_MAX systems per call.  A system whose sizes are out of
 * range (n > 1024, nrhs > 16, lda or ldb < max(1, n)) gets INFO = -1 and is
 * left untouched; every other system's A, IPIV, B and INFO are exactly those
 * of sla_gesv_batched.  Uses the handle's stream, panel width and grid
 * ("batched_w", "batched_wgs_per_cu"; "batched_graph" is ignored) and only
 * enqueues the work: no synchronisation, so the call can be captured.  Calls
 * that overlap on one GPU share the array and race: order them on one
 * stream.  The handle's device must be current. */
#define SLA_BATCHED_NOALLOC_MAX 65536
int sla_gesv_batched_noalloc(sla_handle_t h, int prec, int batch, const int *dn,
                             const int *dnrhs, void *const *dA, const int *dlda,
                             int *const *dipiv, void *const *dB, const int *dldb, int *dinfo);
/* Autotuner: times every panel width and workgroups-per-CU on synthetic
 * systems of these sizes (its own device memory), keeps the fastest in the
 * handle's knobs and returns it (*ms: the batch's time in milliseconds). */
int sla_gesv_batched_tune(sla_handle_t h, int prec, int batch, const int *n, const int *nrhs,
                          int *w, int *wgs_per_cu, double *ms);

/* ------------------------------------------------------------------------ */
/* Part 5: learned precision selector (no counterpart in the reference).      */
/* ------------------------------------------------------------------------ */
/* A network 12 -> 64 -> 64 -> 7 (tanh, softmax; FP32 weights) picks the
 * factorization precision and the refinement of sla_dxgesv -- class 0 FP8 +
 * classical IR, 1 FP8 + GMRES-IR, 2 SLA_PREC_BF16X3 + IR, 3 BF16X3 + GMRES-IR,
 * 4 SLA_PREC_BF16X6 (FP32-accurate factors) + IR, 5 BF16X6 + GMRES-IR, 6 plain
 * FP64 LU -- from 12
 * features of A computed on the GPU (norms, row / column scaling spread,
 * sampled pivot growth and a PDGECON-style condition estimate of sampled
 * principal submatrices, and the stable-rank fraction ||A||_F^2 / (n
 * ||A||_2^2) from two power steps).  Parameters: W1[64][12], b1, W2[64][64],
 * b2, W3[7][64], b3, then the input normalisation mu[12], sigma[12]:
 * SLA_MLP_NPARAM floats. */
#define SLA_SELECT_CLASSES 7
#define SLA_MLP_NPARAM 5471
/* The 12 features of an n x n FP64 device matrix into feat (host float[12]). */
int sla_selector_features(sla_handle_t h, int n, const double *dA, int lda, float *feat);
/* Mean cost-weighted cross-entropy over `batch` samples (dX device
 * float[batch][12] raw features, dy device int labels 0..6, dcost device
 * float[7] class weights or NULL) and its gradient with respect to the
 * trained parameters (the first SLA_MLP_NPARAM - 24 floats of dgrad). */
int sla_mlp_grad(sla_handle_t h, int batch, const float *dX, const int *dy, const float *dcost,
                 const float *dparams, float *dgrad, float *loss);
/* Full-batch Adam (beta 0.9 / 0.999) for `epochs` steps on dparams in place;
 * loss_history: host float[epochs] or NULL. */
int sla_mlp_train(sla_handle_t h, int batch, const float *dX, const int *dy, const float *dcost,
                  float *dparams, int epochs, float lr, float *loss_history);
/* Classes (device int[batch], or NULL) and probabilities (device
 * float[batch][7], or NULL). */
int sla_mlp_predict(sla_handle_t h, int batch, const float *dX, const float *dparams, int *dpred,
                    float *dprob);
/* Install the selector's parameters (host float[SLA_MLP_NPARAM]). */
int sla_selector_set_model(sla_handle_t h, const float *params);
/* sla_dxgesv with the precision and refinement the selector picks (*choice,
 * 0..6) from the features of dA; sla_dxgesv's FP64 fallback stays the judge
 * of the result (a refinement that does not reach FP64 backward error is
 * redone by the FP64 factorization). */
int sla_dxgesv_auto(sla_handle_t h, int n, int nrhs, double *dA, int lda, int *dipiv,
                    const double *dB, int ldb, double *dX, int ldx, int *choice, int *iter,
                    int *info);

/* The low-precision trailing-update kernel alone: C -= A * B (FP32
 * column-major operands, products in `prec`, FP32 accumulation). */
int sla_gemm_lowp_sub(sla_handle_t h, int prec, int m, int n, int k, const float *dA, int lda,
                      const float *dB, int ldb, float *dC, int ldc);

#ifdef __cplusplus
}
#endif

#endif /* SCALAPACK_AMD_H */
