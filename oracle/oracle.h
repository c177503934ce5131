/*
 * oracle.h -- CPU oracle for the dense LU solve path (TEST INFRASTRUCTURE).
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
 * library, and only as a checker or a timed CPU baseline.  The product library
 * (scalapack_amd) never links or calls it.
 *
 * Every routine restates the reference's 1 x 1 BLACS-grid behaviour:
 *   oracle_pdgetrf  SRC/pdgetrf.f:217-295     (blocked right-looking LU)
 *   oracle_pdgetf2  SRC/pdgetf2.f:206-237     (unblocked panel, PDAMAX/PDSWAP/
 *                                              PDSCAL(1/GMAX)/PDGER)
 *   oracle_pdlaswp  SRC/pdlaswp.f:163-172     (forward row interchanges)
 *   oracle_pdgetrs  SRC/pdgetrs.f:244-286     (PDLAPIV + two PDTRSM)
 *   oracle_pdgesv   SRC/pdgesv.f:229-236
 *   oracle_matgen   TESTING/traditional/LIN/pdmatgen.f:448-510 and
 *                   pmatgeninc.f (the 31-bit LCG, closed-form jump-ahead)
 *   oracle_sresid   TESTING/traditional/LIN/pdlaschk.f:185-187,274-299
 *   oracle_fresid   TESTING/traditional/LIN/pdlafchk.f:199-226 with the
 *                   P*L*U rebuild of pdgetrrv.f:210-289
 *   oracle_pdlange  SRC/pdlange.f:196-330          (aux_oracle.c)
 *   oracle_pdgecon  SRC/pdgecon.f:287-403 with SRC/pdlacon.f
 *   oracle_pdgerfs  SRC/pdgerfs.f:440-675 (PDAGEMV as PTZBLAS/dagemv.f)
 *   oracle_pdgetrs_t SRC/pdgetrs.f:273-284       (TRANS = 'T')
 *   oracle_pdgeequ  SRC/pdgeequ.f; oracle_pdlaqge SRC/pdlaqge.f:205-262
 *   oracle_pdgesvx  SRC/pdgesvx.f:624-737 (FACT = 'E' or 'N')
 * Arithmetic goes through the host Fortran BLAS (dgemm_, dtrsm_, idamax_,
 * dswap_, dscal_, dger_) in the order the reference issues those calls, with
 * PDTRSM's 32-row blocking (PBLAS/SRC/PTOOLS/PB_CptrsmAB.c:301-412,
 * PBLAS/SRC/pilaenv.f:75) and PDGEMM's single local dgemm_
 * (PBLAS/SRC/PTOOLS/PB_CpgemmAB.c:333-352).
 *
 * Indices follow the Fortran interface: IA, JA, IB, JB are 1-based global
 * indices, IPIV holds 1-based global row numbers, DESC is the 9-entry
 * descriptor (DTYPE, CTXT, M, N, MB, NB, RSRC, CSRC, LLD).
 */
#ifndef SCALAPACK_AMD_ORACLE_H
#define SCALAPACK_AMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int oracle_pdgetf2(int m, int n, double *a, int ia, int ja, const int *desca,
                   int *ipiv, double *gap);
void oracle_pdlaswp(char direc, int n, double *a, int ia, int ja,
                    const int *desca, int k1, int k2, const int *ipiv);
int oracle_pdgetrf(int m, int n, double *a, int ia, int ja, const int *desca,
                   int *ipiv, double *gap);
int oracle_pdgetrs(char trans, int n, int nrhs, const double *a, int ia, int ja,
                   const int *desca, const int *ipiv, double *b, int ib, int jb,
                   const int *descb);
int oracle_pdgesv(int n, int nrhs, double *a, int ia, int ja, const int *desca,
                  int *ipiv, double *b, int ib, int jb, const int *descb);

/* Entries A(i0+i, j0+j), 0 <= i < mloc, 0 <= j < nloc, of the global M x N
 * PDMATGEN matrix with seed `seed`, written column-major into out (ld ldo). */
void oracle_matgen(int M, int N, int seed, int i0, int j0, int mloc, int nloc,
                   double *out, int ldo);
/* The LCG state after k steps from seed (X_k, 0 <= X_k < 2^31). */
unsigned oracle_lcg_state(int seed, long long k);

double oracle_pdlange(char norm, int m, int n, const double *a, int ia, int ja,
                      const int *desca);
/* Returns RCOND for the L\U factors at sub(A). */
double oracle_pdgecon(char norm, int n, const double *a, int ia, int ja, const int *desca,
                      double anorm);
int oracle_pdgetrs_t(int n, int nrhs, const double *a, int ia, int ja, const int *desca,
                     const int *ipiv, double *b, int ib, int jb, const int *descb);
/* IA = JA = IAF = JAF = IB = JB = IX = JX = 1; IPIV 1-based. */
int oracle_pdgerfs(char trans, int n, int nrhs, const double *a, int lda, const double *af,
                   int ldaf, const int *ipiv, const double *b, int ldb, double *x, int ldx,
                   double *ferr, double *berr);

/* PDGEEQU / PDLAQGE / PDGESVX (aux_oracle.c), IA = JA = 1. */
int oracle_pdgeequ(int m, int n, const double *a, int lda, double *r, double *c, double *rowcnd,
                   double *colcnd, double *amax);
char oracle_pdlaqge(int m, int n, double *a, int lda, const double *r, const double *c,
                    double rowcnd, double colcnd, double amax);
int oracle_pdgesvx(char fact, char trans, int n, int nrhs, int nb, double *a, double *af,
                   int *ipiv, char *equed, double *r, double *c, double *b, double *x,
                   double *rcond, double *ferr, double *berr);

double oracle_norm_inf(int m, int n, const double *a, int lda);
double oracle_sresid(int n, int nrhs, const double *a, int lda, const double *x,
                     int ldx, const double *b, int ldb);
double oracle_fresid(int m, int n, const double *a0, int lda0, const double *lu,
                     int ldlu, const int *ipiv);

#ifdef __cplusplus
}
#endif
#endif
