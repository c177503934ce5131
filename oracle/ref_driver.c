/*
 * ref_driver.c -- runs the reference ScaLAPACK LU path (PDGETRF + PDGETRS,
 * compiled from the reference sources by oracle/Makefile into oracle/_ref/)
 * on a P x Q BLACS grid and writes what the tests and the CPU baseline need.
 *
 * TEST INFRASTRUCTURE ONLY: this program is a checker and a baseline timer.
 * It is never linked into, loaded by or called from the product library.
 *
 * It mirrors the reference LU tester's sequence
 * (TESTING/traditional/LIN/pdludriver.f:331-335 PDMATGEN of A with IASEED,
 *  :374-383 BLACS_BARRIER + timed PDGETRF, :665 PDMATGEN of B with IBSEED,
 *  :704-713 timed PDGETRS) and its residual
 * (TESTING/traditional/LIN/pdlaschk.f:185-187,274-299:
 *  max_j ||A x_j - b_j||_inf / (||A||_inf ||x_j||_inf eps N), eps = 2^-53).
 *
 * usage: ref_lu M N NB NRHS P Q SEEDA BMODE OUT [--lu] [--nosolve]
 *        ref_lu aux N NB NRHS SEEDA SEEDB OUT      (1 x 1 grid, see aux())
 *        ref_lu svx N NB NRHS OUT                  (1 x 1 grid, see svx())
 *        ref_lu io M N MB FILE IN OUT              (1 x 1 grid, see io())
 *        ref_lu argcheck ROUTINE ARGS...           (see argcheck())
 *   BMODE = lcg:<seed>  (PDMATGEN of B, as the tester)
 *         | ones        (b = A * ones(N), summed left to right in double)
 *         | file:<path> (N x NRHS column-major doubles)
 * OUT is written by rank 0:
 *   int32  magic 0x4C55524E, M, N, NB, NRHS, P, Q, info_getrf, info_getrs, flags
 *   double t_getrf, t_getrs, sresid, anorm_inf
 *   int32  ipiv[min(M,N)]              (global, 1-based, as PDGETRF returns)
 *   double x[N*NRHS]                   (if a solve ran)
 *   double lu[M*N]                     (if --lu)
 */
#include <mpi.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

extern void blacs_pinfo_(int *, int *);
extern void blacs_get_(int *, int *, int *);
extern void blacs_gridinit_(int *, const char *, int *, int *, size_t);
extern void blacs_gridinfo_(int *, int *, int *, int *, int *);
extern void blacs_barrier_(int *, const char *, size_t);
extern void blacs_gridexit_(int *);
extern void blacs_exit_(int *);
extern int numroc_(int *, int *, int *, int *, int *);
extern void descinit_(int *, int *, int *, int *, int *, int *, int *, int *,
                      int *, int *);
extern void pdmatgen_(int *, const char *, const char *, int *, int *, int *,
                      int *, double *, int *, int *, int *, int *, int *, int *,
                      int *, int *, int *, int *, int *, int *, size_t, size_t);
extern void pdgetrf_(int *, int *, double *, int *, int *, int *, int *, int *);
extern void pdgetrs_(const char *, int *, int *, double *, int *, int *, int *,
                     int *, double *, int *, int *, int *, int *, size_t);

static int g2p(int g, int nb, int np) { return (g / nb) % np; }
static int g2l(int g, int nb, int np) { return (g / (nb * np)) * nb + g % nb; }

/* Generate the whole global M x N PDMATGEN matrix on this rank by calling the
 * reference generator with a 1 x 1 view (the stream is a function of the
 * global index only, pdmatgen.f:448-510). */
static void gen_global(int ictxt, int m, int n, int nb, int seed, double *g)
{
    int zero = 0, one = 1, ldg = m > 1 ? m : 1;
    pdmatgen_(&ictxt, "N", "N", &m, &n, &nb, &nb, g, &ldg, &zero, &zero, &seed,
              &zero, &m, &zero, &n, &zero, &zero, &one, &one, 1, 1);
}

extern void pdgesv_(int *, int *, double *, int *, int *, int *, int *, double *, int *,
                    int *, int *, int *);

/* argcheck ROUTINE v1 v2 ... : one call of PDGETRF / PDGETRS / PDGESV with the
 * given scalar arguments and descriptors on a 1 x 1 grid; prints INFO.  A
 * descriptor CTXT entry of -999 is replaced by an invalid context, any other
 * value by the valid one.
 *   pdgetrf M N IA JA DESCA(9)
 *   pdgetrs TRANS N NRHS IA JA DESCA(9) IB JB DESCB(9)
 *   pdgesv  N NRHS IA JA DESCA(9) IB JB DESCB(9)                             */
static int argcheck(int argc, char **argv)
{
    int ictxt, neg1 = -1, zero = 0, one = 1;
    blacs_get_(&neg1, &zero, &ictxt);
    blacs_gridinit_(&ictxt, "Row-major", &one, &one, 9);
    const char *r = argv[2];
    int v[64], nv = 0;
    for (int i = 3; i < argc && nv < 64; i++) v[nv++] = atoi(argv[i]);
    static double a[4096], b[4096];
    static int ipiv[512];
    int info = 12345;
    int *da, *db;
    for (int i = 0; i < 512; i++) ipiv[i] = i + 1;   /* identity pivots */
#define FIXCTX(d) do { if ((d)[1] == -999) (d)[1] = ictxt + 77; else (d)[1] = ictxt; } while (0)
    /* identity A (with the descriptor's LLD) so that legal calls factor and
     * solve a regular system */
#define EYE(d) do { int ld_ = (d)[8] > 0 ? (d)[8] : 1; \
        for (int i_ = 0; i_ < 32 && i_ * (ld_ + 1) < 4096; i_++) a[i_ * (ld_ + 1)] = 1.0; } while (0)
    if (!strcmp(r, "pdgetrf")) {
        da = v + 4; FIXCTX(da); EYE(da);
        pdgetrf_(&v[0], &v[1], a, &v[2], &v[3], da, ipiv, &info);
    } else if (!strcmp(r, "pdgetrs")) {
        char t = argv[3][0];
        int *w = v + 1;   /* N NRHS IA JA DESCA IB JB DESCB */
        da = w + 4; db = w + 15; FIXCTX(da); FIXCTX(db); EYE(da);
        pdgetrs_(&t, &w[0], &w[1], a, &w[2], &w[3], da, ipiv, b, &w[13], &w[14], db, &info, 1);
    } else if (!strcmp(r, "pdgesv")) {
        da = v + 4; db = v + 15; FIXCTX(da); FIXCTX(db); EYE(da);
        pdgesv_(&v[0], &v[1], a, &v[2], &v[3], da, ipiv, b, &v[13], &v[14], db, &info);
    }
    printf("INFO %d\n", info);
    blacs_gridexit_(&ictxt);
    int cont = 0;
    blacs_exit_(&cont);
    return 0;
}

extern double pdlange_(const char *, int *, int *, double *, int *, int *, int *, double *,
                       size_t);
extern void pdgecon_(const char *, int *, double *, int *, int *, int *, double *, double *,
                     double *, int *, int *, int *, int *, size_t);
extern void pdgerfs_(const char *, int *, int *, double *, int *, int *, int *, double *, int *,
                     int *, int *, int *, double *, int *, int *, int *, double *, int *, int *,
                     int *, double *, double *, double *, int *, int *, int *, int *, size_t);

/* aux N NB NRHS SEEDA SEEDB OUT: the routines beside the solve on a 1 x 1
 * grid, A = PDMATGEN(SEEDA) N x N, B = PDMATGEN(SEEDB) N x NRHS.
 *   PDLANGE 'M','1','I','F' of A and of the submatrix A(2:N-2, 3:N)
 *   PDGETRF of A, PDGECON '1' and 'I' with the matching PDLANGE norms
 *   PDGETRS 'T' of B
 *   PDGERFS 'N' from the PDGETRS 'N' solution and from a perturbed start,
 *   PDGERFS 'T' from a perturbed start of the 'T' solution
 * OUT: int32 magic 0x41555852, N, NB, NRHS, info_getrf, info_gecon1,
 *      info_geconI, info_getrsT, info_rfs[3]
 *      double norms[4], subnorms[4], rcond1, rcondI
 *      int32 ipiv[N]
 *      double xT[N*NRHS], then per refinement case c = 0, 1, 2:
 *      x0[N*NRHS] (start), x[N*NRHS] (refined), ferr[NRHS], berr[NRHS] */
static int aux(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: ref_lu aux N NB NRHS SEEDA SEEDB OUT\n");
        return 2;
    }
    int N = atoi(argv[2]), NB = atoi(argv[3]), NRHS = atoi(argv[4]);
    int seedA = atoi(argv[5]), seedB = atoi(argv[6]);
    const char *out = argv[7];
    int ictxt, neg1 = -1, zero = 0, one = 1, info = 0;
    blacs_get_(&neg1, &zero, &ictxt);
    blacs_gridinit_(&ictxt, "Row-major", &one, &one, 9);
    int lld = N > 1 ? N : 1;
    int desca[9], descb[9];
    descinit_(desca, &N, &N, &NB, &NB, &zero, &zero, &ictxt, &lld, &info);
    descinit_(descb, &N, &NRHS, &NB, &NB, &zero, &zero, &ictxt, &lld, &info);
    size_t nn = (size_t)N * N, nr = (size_t)N * NRHS;
    double *A = malloc(nn * 8), *AF = malloc(nn * 8), *B = malloc(nr * 8);
    double *X = malloc(nr * 8), *XT = malloc(nr * 8), *X0 = malloc(3 * nr * 8),
           *XR = malloc(3 * nr * 8);
    double *ferr = calloc(3 * NRHS, 8), *berr = calloc(3 * NRHS, 8);
    int *ipiv = calloc((size_t)N + NB, sizeof(int));
    gen_global(ictxt, N, N, NB, seedA, A);
    gen_global(ictxt, N, NRHS, NB, seedB, B);

    const char nrm[4] = {'M', '1', 'I', 'F'};
    double norms[4], subnorms[4];
    double *lw = malloc(((size_t)N + 64) * 8);
    for (int k = 0; k < 4; k++) norms[k] = pdlange_(&nrm[k], &N, &N, A, &one, &one, desca, lw, 1);
    {
        int ms = N - 3, ns = N - 2, ia = 2, ja = 3;
        for (int k = 0; k < 4; k++)
            subnorms[k] = pdlange_(&nrm[k], &ms, &ns, A, &ia, &ja, desca, lw, 1);
    }
    memcpy(AF, A, nn * 8);
    int info_f = 0;
    pdgetrf_(&N, &N, AF, &one, &one, desca, ipiv, &info_f);

    /* PDGECON with workspace from its own query */
    double rcond[2] = {0.0, 0.0};
    int info_c[2] = {0, 0};
    for (int k = 0; k < 2; k++) {
        const char cn = k == 0 ? '1' : 'I';
        double anorm = k == 0 ? norms[1] : norms[2], wq;
        int iq, lq = -1, inf = 0;
        pdgecon_(&cn, &N, AF, &one, &one, desca, &anorm, &rcond[k], &wq, &lq, &iq, &lq, &inf, 1);
        int lwork = (int)wq, liwork = iq;
        double *work = malloc((size_t)lwork * 8);
        int *iwork = malloc((size_t)liwork * 4);
        pdgecon_(&cn, &N, AF, &one, &one, desca, &anorm, &rcond[k], work, &lwork, iwork, &liwork,
                 &info_c[k], 1);
        free(work);
        free(iwork);
    }

    /* PDGETRS 'N' and 'T' */
    int info_t = 0, info_n = 0;
    memcpy(XT, B, nr * 8);
    pdgetrs_("T", &N, &NRHS, AF, &one, &one, desca, ipiv, XT, &one, &one, descb, &info_t, 1);
    memcpy(X, B, nr * 8);
    pdgetrs_("N", &N, &NRHS, AF, &one, &one, desca, ipiv, X, &one, &one, descb, &info_n, 1);

    /* PDGERFS: case 0 'N' from X, case 1 'N' from X perturbed, case 2 'T'
     * from XT perturbed; perturbation x_i (1 + 1e-7 ((i mod 7) - 3)). */
    int info_r[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) {
        double *x0 = X0 + (size_t)c * nr, *xr = XR + (size_t)c * nr;
        const double *src = c == 2 ? XT : X;
        for (size_t i = 0; i < nr; i++) {
            const int row = (int)(i % (size_t)N);
            x0[i] = c == 0 ? src[i] : src[i] * (1.0 + 1e-7 * (double)((row % 7) - 3));
        }
        memcpy(xr, x0, nr * 8);
        const char tr = c == 2 ? 'T' : 'N';
        double wq;
        int iq, lq = -1, inf = 0;
        pdgerfs_(&tr, &N, &NRHS, A, &one, &one, desca, AF, &one, &one, desca, ipiv, B, &one, &one,
                 descb, xr, &one, &one, descb, ferr + c * NRHS, berr + c * NRHS, &wq, &lq, &iq,
                 &lq, &inf, 1);
        int lwork = (int)wq, liwork = iq;
        double *work = malloc((size_t)lwork * 8);
        int *iwork = malloc((size_t)liwork * 4);
        pdgerfs_(&tr, &N, &NRHS, A, &one, &one, desca, AF, &one, &one, desca, ipiv, B, &one, &one,
                 descb, xr, &one, &one, descb, ferr + c * NRHS, berr + c * NRHS, work, &lwork,
                 iwork, &liwork, &info_r[c], 1);
        free(work);
        free(iwork);
    }

    FILE *f = fopen(out, "wb");
    int32_t hdr[11] = {0x41555852, N, NB, NRHS, info_f, info_c[0], info_c[1], info_t,
                       info_r[0], info_r[1], info_r[2]};
    fwrite(hdr, sizeof hdr, 1, f);
    fwrite(norms, 8, 4, f);
    fwrite(subnorms, 8, 4, f);
    fwrite(rcond, 8, 2, f);
    fwrite(ipiv, 4, (size_t)N, f);
    fwrite(XT, 8, nr, f);
    for (int c = 0; c < 3; c++) {
        fwrite(X0 + (size_t)c * nr, 8, nr, f);
        fwrite(XR + (size_t)c * nr, 8, nr, f);
        fwrite(ferr + c * NRHS, 8, (size_t)NRHS, f);
        fwrite(berr + c * NRHS, 8, (size_t)NRHS, f);
    }
    fclose(f);
    printf("{\"N\": %d, \"NB\": %d, \"NRHS\": %d, \"rcond1\": %.17g, \"rcondI\": %.17g, "
           "\"info\": [%d, %d, %d, %d, %d, %d, %d]}\n",
           N, NB, NRHS, rcond[0], rcond[1], info_f, info_c[0], info_c[1], info_t, info_r[0],
           info_r[1], info_r[2]);
    free(A); free(AF); free(B); free(X); free(XT); free(X0); free(XR); free(ferr); free(berr);
    free(ipiv); free(lw);
    blacs_gridexit_(&ictxt);
    int cont = 0;
    blacs_exit_(&cont);
    return 0;
}

extern void pdgeequ_(int *, int *, double *, int *, int *, int *, double *, double *, double *,
                     double *, double *, int *);
extern void pdlaqge_(int *, int *, double *, int *, int *, int *, double *, double *, double *,
                     double *, double *, char *, size_t);
extern void pdgesvx_(const char *, const char *, int *, int *, double *, int *, int *, int *,
                     double *, int *, int *, int *, int *, char *, double *, double *, double *,
                     int *, int *, int *, double *, int *, int *, int *, double *, double *,
                     double *, double *, int *, int *, int *, int *, size_t, size_t, size_t);

/* svx N NB NRHS OUT: equilibration and the expert driver on a 1 x 1 grid.
 * A = PDMATGEN(N, N, 100) with A(i,j) scaled by 2^(3 (i mod 7) - 4 (j mod 5))
 * (0-based, exact), B = PDMATGEN(N, NRHS, 200).
 *   PDGEEQU of A; PDLAQGE of A with its output
 *   PDGESVX cases: 0 FACT='E' TRANS='N', 1 'E' 'T', 2 'N' 'N'
 * OUT: int32 magic 0x53565852, N, NB, NRHS, info_equ, equed_laqge,
 *      per case: info, equed
 *      double rowcnd, colcnd, amax, R[N], C[N], A_laqge[N*N]
 *      per case: rcond, R[N], C[N], X[N*NRHS], FERR[NRHS], BERR[NRHS],
 *                B_out[N*NRHS] */
static int svx(int argc, char **argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: ref_lu svx N NB NRHS OUT\n");
        return 2;
    }
    int N = atoi(argv[2]), NB = atoi(argv[3]), NRHS = atoi(argv[4]);
    const char *out = argv[5];
    int ictxt, neg1 = -1, zero = 0, one = 1, info = 0;
    blacs_get_(&neg1, &zero, &ictxt);
    blacs_gridinit_(&ictxt, "Row-major", &one, &one, 9);
    int lld = N > 1 ? N : 1;
    int desca[9], descb[9];
    descinit_(desca, &N, &N, &NB, &NB, &zero, &zero, &ictxt, &lld, &info);
    descinit_(descb, &N, &NRHS, &NB, &NB, &zero, &zero, &ictxt, &lld, &info);
    size_t nn = (size_t)N * N, nr = (size_t)N * NRHS;
    double *A0 = malloc(nn * 8), *A = malloc(nn * 8), *AF = malloc(nn * 8);
    double *B0 = malloc(nr * 8), *B = malloc(nr * 8), *X = malloc(nr * 8);
    double *R = malloc((size_t)N * 8), *C = malloc((size_t)N * 8);
    int *ipiv = calloc((size_t)N + NB, sizeof(int));
    gen_global(ictxt, N, N, NB, 100, A0);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) A0[i + (size_t)j * N] = ldexp(A0[i + (size_t)j * N], 3 * (i % 7) - 4 * (j % 5));
    gen_global(ictxt, N, NRHS, NB, 200, B0);

    double rowcnd = 0, colcnd = 0, amax = 0;
    int info_eq = 0;
    memcpy(A, A0, nn * 8);
    pdgeequ_(&N, &N, A, &one, &one, desca, R, C, &rowcnd, &colcnd, &amax, &info_eq);
    char equed_l = 'X';
    pdlaqge_(&N, &N, A, &one, &one, desca, R, C, &rowcnd, &colcnd, &amax, &equed_l, 1);

    FILE *f = fopen(out, "wb");
    int32_t hdr[6] = {0x53565852, N, NB, NRHS, info_eq, (int)equed_l};
    fwrite(hdr, sizeof hdr, 1, f);
    /* the cases first into memory, so the header can carry their codes */
    const char facts[3] = {'E', 'E', 'N'}, transs[3] = {'N', 'T', 'N'};
    int32_t cinfo[6];
    double *blob[3];
    size_t blen = 1 + 2 * (size_t)N + nr + 2 * (size_t)NRHS + nr;
    for (int c = 0; c < 3; c++) {
        double *Rc = malloc((size_t)N * 8), *Cc = malloc((size_t)N * 8);
        double *ferr = calloc((size_t)NRHS, 8), *berr = calloc((size_t)NRHS, 8);
        double rcond = -1.0;
        char equed = 'N';
        memcpy(A, A0, nn * 8);
        memcpy(B, B0, nr * 8);
        memset(X, 0, nr * 8);
        double wq;
        int iq, lq = -1, inf = 0;
        pdgesvx_(&facts[c], &transs[c], &N, &NRHS, A, &one, &one, desca, AF, &one, &one, desca, ipiv,
                 &equed, Rc, Cc, B, &one, &one, descb, X, &one, &one, descb, &rcond, ferr, berr,
                 &wq, &lq, &iq, &lq, &inf, 1, 1, 1);
        int lwork = (int)wq, liwork = iq;
        double *work = malloc((size_t)lwork * 8);
        int *iwork = malloc((size_t)liwork * 4);
        pdgesvx_(&facts[c], &transs[c], &N, &NRHS, A, &one, &one, desca, AF, &one, &one, desca, ipiv,
                 &equed, Rc, Cc, B, &one, &one, descb, X, &one, &one, descb, &rcond, ferr, berr,
                 work, &lwork, iwork, &liwork, &inf, 1, 1, 1);
        cinfo[2 * c] = inf;
        cinfo[2 * c + 1] = (int)equed;
        double *bl = malloc(blen * 8), *q = bl;
        *q++ = rcond;
        memcpy(q, Rc, (size_t)N * 8); q += N;
        memcpy(q, Cc, (size_t)N * 8); q += N;
        memcpy(q, X, nr * 8); q += nr;
        memcpy(q, ferr, (size_t)NRHS * 8); q += NRHS;
        memcpy(q, berr, (size_t)NRHS * 8); q += NRHS;
        memcpy(q, B, nr * 8);
        blob[c] = bl;
        free(work); free(iwork); free(Rc); free(Cc); free(ferr); free(berr);
    }
    fwrite(cinfo, sizeof cinfo, 1, f);
    double eqd[3] = {rowcnd, colcnd, amax};
    fwrite(eqd, 8, 3, f);
    fwrite(R, 8, (size_t)N, f);
    fwrite(C, 8, (size_t)N, f);
    fclose(f);
    /* A after PDLAQGE: recompute on a fresh copy (A was reused by the cases) */
    memcpy(A, A0, nn * 8);
    {
        double r2, c2, am2;
        int i2;
        pdgeequ_(&N, &N, A, &one, &one, desca, R, C, &r2, &c2, &am2, &i2);
        char e2;
        pdlaqge_(&N, &N, A, &one, &one, desca, R, C, &r2, &c2, &am2, &e2, 1);
    }
    f = fopen(out, "ab");
    fwrite(A, 8, nn, f);
    for (int c = 0; c < 3; c++) {
        fwrite(blob[c], 8, blen, f);
        free(blob[c]);
    }
    fclose(f);
    printf("{\"N\": %d, \"info_equ\": %d, \"equed\": \"%c\", \"cases\": [[%d, \"%c\"], [%d, \"%c\"], "
           "[%d, \"%c\"]]}\n", N, info_eq, equed_l, cinfo[0], (char)cinfo[1], cinfo[2],
           (char)cinfo[3], cinfo[4], (char)cinfo[5]);
    free(A0); free(A); free(AF); free(B0); free(B); free(X); free(R); free(C); free(ipiv);
    blacs_gridexit_(&ictxt);
    int cont = 0;
    blacs_exit_(&cont);
    return 0;
}

extern void pdlawrite_(const char *, int *, int *, double *, int *, int *, int *, int *, int *,
                       double *, size_t);
extern void pdlaread_(const char *, double *, int *, int *, int *, double *, size_t);

/* io M N MB FILE IN OUT: PDLAWRITE of the M x N matrix PDMATGEN(M, N, 100)
 * with a few special values (0, 1, -0.5, 1e-300, -3e300, 123456789) at
 * fixed places, as sub(A) = A(1:M, 1:N) of an (M+2) x (N+1) array, to FILE;
 * then PDLAREAD of the file IN into an M_A x N_A = (M+3) x (N+2) array
 * (zero filled) written to OUT as raw doubles (column-major, LLD = M+3). */
static int io(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: ref_lu io M N MB FILE IN OUT\n");
        return 2;
    }
    int M = atoi(argv[2]), N = atoi(argv[3]), MB = atoi(argv[4]);
    const char *file = argv[5], *in = argv[6], *out = argv[7];
    int ictxt, neg1 = -1, zero = 0, one = 1, info = 0;
    blacs_get_(&neg1, &zero, &ictxt);
    blacs_gridinit_(&ictxt, "Row-major", &one, &one, 9);
    int Mg = M + 2, Ng = N + 1, lld = Mg;
    int desca[9];
    descinit_(desca, &Mg, &Ng, &MB, &MB, &zero, &zero, &ictxt, &lld, &info);
    double *A = calloc((size_t)Mg * Ng, 8), *work = calloc((size_t)MB + 1, 8);
    double *G = malloc((size_t)M * N * 8);
    gen_global(ictxt, M, N, MB, 100, G);
    const double special[6] = {0.0, 1.0, -0.5, 1e-300, -3e300, 123456789.0};
    for (int k = 0; k < 6 && k < M * N; k++) G[(size_t)(k * 7) % ((size_t)M * N)] = special[k];
    for (int j = 0; j < N; j++)
        for (int i = 0; i < M; i++) A[i + (size_t)j * lld] = G[i + (size_t)j * M];
    pdlawrite_(file, &M, &N, A, &one, &one, desca, &zero, &zero, work, strlen(file));
    int Mr = M + 3, Nr = N + 2, lldr = Mr, descr[9];
    descinit_(descr, &Mr, &Nr, &MB, &MB, &zero, &zero, &ictxt, &lldr, &info);
    double *Ar = calloc((size_t)Mr * Nr, 8), *wr = calloc((size_t)MB + 1, 8);
    pdlaread_(in, Ar, descr, &zero, &zero, wr, strlen(in));
    FILE *f = fopen(out, "wb");
    fwrite(Ar, 8, (size_t)Mr * Nr, f);
    fclose(f);
    free(A); free(work); free(G); free(Ar); free(wr);
    blacs_gridexit_(&ictxt);
    int cont = 0;
    blacs_exit_(&cont);
    return 0;
}

int main(int argc, char **argv)
{
    int iam, nprocs;
    blacs_pinfo_(&iam, &nprocs);
    if (argc > 2 && !strcmp(argv[1], "argcheck")) return argcheck(argc, argv);
    if (argc > 1 && !strcmp(argv[1], "aux")) return aux(argc, argv);
    if (argc > 1 && !strcmp(argv[1], "svx")) return svx(argc, argv);
    if (argc > 1 && !strcmp(argv[1], "io")) return io(argc, argv);
    if (argc < 10) {
        if (iam == 0)
            fprintf(stderr, "usage: ref_lu M N NB NRHS P Q SEEDA BMODE OUT [--lu] [--nosolve]\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    int M = atoi(argv[1]), N = atoi(argv[2]), NB = atoi(argv[3]);
    int NRHS = atoi(argv[4]), P = atoi(argv[5]), Q = atoi(argv[6]);
    int seedA = atoi(argv[7]);
    const char *bmode = argv[8], *out = argv[9];
    int want_lu = 0, solve = (M == N);
    for (int i = 10; i < argc; i++) {
        if (!strcmp(argv[i], "--lu")) want_lu = 1;
        if (!strcmp(argv[i], "--nosolve")) solve = 0;
    }
    if (P * Q != nprocs) {
        if (iam == 0) fprintf(stderr, "grid %dx%d needs %d ranks, have %d\n", P, Q, P * Q, nprocs);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    int ictxt, neg1 = -1, zero = 0, one = 1, info = 0;
    blacs_get_(&neg1, &zero, &ictxt);
    blacs_gridinit_(&ictxt, "Row-major", &P, &Q, 9);
    int nprow, npcol, myrow, mycol;
    blacs_gridinfo_(&ictxt, &nprow, &npcol, &myrow, &mycol);

    int mp = numroc_(&M, &NB, &myrow, &zero, &nprow);
    int nq = numroc_(&N, &NB, &mycol, &zero, &npcol);
    int lld = mp > 1 ? mp : 1;
    int desca[9], descb[9];
    descinit_(desca, &M, &N, &NB, &NB, &zero, &zero, &ictxt, &lld, &info);
    int nrq = numroc_(&NRHS, &NB, &mycol, &zero, &npcol);
    int npb = numroc_(&N, &NB, &myrow, &zero, &nprow);
    int lldb = npb > 1 ? npb : 1;
    descinit_(descb, &N, &NRHS, &NB, &NB, &zero, &zero, &ictxt, &lldb, &info);

    double *A = calloc((size_t)lld * (nq > 0 ? nq : 1), sizeof(double));
    double *B = calloc((size_t)lldb * (nrq > 0 ? nrq : 1), sizeof(double));
    int *ipiv = calloc((size_t)mp + NB, sizeof(int));

    /* A: PDMATGEN with the tester's argument list (pdludriver.f:331-335). */
    pdmatgen_(&ictxt, "N", "N", &desca[2], &desca[3], &desca[4], &desca[5], A,
              &desca[8], &desca[6], &desca[7], &seedA, &zero, &mp, &zero, &nq,
              &myrow, &mycol, &nprow, &npcol, 1, 1);

    /* Global copies needed by B = A*ones and the residual. */
    double *Ag = NULL, *Bg = NULL;
    int need_global = (iam == 0) || !strcmp(bmode, "ones") || !strncmp(bmode, "file:", 5);
    if (solve && need_global) {
        Bg = calloc((size_t)N * NRHS, sizeof(double));
        if (!strncmp(bmode, "lcg:", 4)) {
            int sb = atoi(bmode + 4);
            gen_global(ictxt, N, NRHS, NB, sb, Bg);
        } else if (!strcmp(bmode, "ones")) {
            Ag = malloc((size_t)M * N * sizeof(double));
            gen_global(ictxt, M, N, NB, seedA, Ag);
            for (int i = 0; i < N; i++) {
                double s = 0.0;
                for (int j = 0; j < N; j++) s += Ag[i + (size_t)j * M];
                for (int r = 0; r < NRHS; r++) Bg[i + (size_t)r * N] = s;
            }
        } else if (!strncmp(bmode, "file:", 5)) {
            FILE *f = fopen(bmode + 5, "rb");
            if (!f || fread(Bg, sizeof(double), (size_t)N * NRHS, f) != (size_t)N * NRHS) {
                fprintf(stderr, "cannot read %s\n", bmode + 5);
                MPI_Abort(MPI_COMM_WORLD, 2);
            }
            fclose(f);
        }
    }
    if (solve) {
        if (!strncmp(bmode, "lcg:", 4)) {
            int sb = atoi(bmode + 4);
            pdmatgen_(&ictxt, "N", "N", &descb[2], &descb[3], &descb[4], &descb[5], B,
                      &descb[8], &descb[6], &descb[7], &sb, &zero, &npb, &zero, &nrq,
                      &myrow, &mycol, &nprow, &npcol, 1, 1);
        } else {
            for (int gj = 0; gj < NRHS; gj++) {
                if (g2p(gj, NB, npcol) != mycol) continue;
                int lj = g2l(gj, NB, npcol);
                for (int gi = 0; gi < N; gi++) {
                    if (g2p(gi, NB, nprow) != myrow) continue;
                    B[g2l(gi, NB, nprow) + (size_t)lj * lldb] = Bg[gi + (size_t)gj * N];
                }
            }
        }
    }

    int info_f = 0, info_s = 0;
    double t0, t_getrf, t_getrs = 0.0;
    blacs_barrier_(&ictxt, "All", 3);
    t0 = MPI_Wtime();
    pdgetrf_(&M, &N, A, &one, &one, desca, ipiv, &info_f);
    t_getrf = MPI_Wtime() - t0;
    MPI_Allreduce(MPI_IN_PLACE, &t_getrf, 1, MPI_DOUBLE, MPI_MAX, MPI_COMM_WORLD);
    if (solve && info_f == 0) {
        blacs_barrier_(&ictxt, "All", 3);
        t0 = MPI_Wtime();
        pdgetrs_("N", &N, &NRHS, A, &one, &one, desca, ipiv, B, &one, &one, descb,
                 &info_s, 1);
        t_getrs = MPI_Wtime() - t0;
        MPI_Allreduce(MPI_IN_PLACE, &t_getrs, 1, MPI_DOUBLE, MPI_MAX, MPI_COMM_WORLD);
    } else {
        solve = 0;
    }

    /* Gather IPIV (process column 0 holds it for its local rows), X and LU. */
    int MN = M < N ? M : N;
    int *ipg = calloc((size_t)MN + 1, sizeof(int));
    for (int gi = 0; gi < MN; gi++)
        if (g2p(gi, NB, nprow) == myrow && mycol == 0)
            ipg[gi] = ipiv[g2l(gi, NB, nprow)];
    MPI_Allreduce(MPI_IN_PLACE, ipg, MN, MPI_INT, MPI_SUM, MPI_COMM_WORLD);

    double *Xg = NULL, *LUg = NULL;
    if (solve) {
        Xg = calloc((size_t)N * NRHS, sizeof(double));
        for (int gj = 0; gj < NRHS; gj++) {
            if (g2p(gj, NB, npcol) != mycol) continue;
            for (int gi = 0; gi < N; gi++)
                if (g2p(gi, NB, nprow) == myrow)
                    Xg[gi + (size_t)gj * N] = B[g2l(gi, NB, nprow) + (size_t)g2l(gj, NB, npcol) * lldb];
        }
        MPI_Allreduce(MPI_IN_PLACE, Xg, N * NRHS, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    }
    if (want_lu) {
        LUg = calloc((size_t)M * N, sizeof(double));
        for (int gj = 0; gj < N; gj++) {
            if (g2p(gj, NB, npcol) != mycol) continue;
            for (int gi = 0; gi < M; gi++)
                if (g2p(gi, NB, nprow) == myrow)
                    LUg[gi + (size_t)gj * M] = A[g2l(gi, NB, nprow) + (size_t)g2l(gj, NB, npcol) * lld];
        }
        if (nprocs > 1)
            MPI_Allreduce(MPI_IN_PLACE, LUg, M * N, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    }

    if (iam == 0) {
        double sresid = -1.0, anorm = 0.0;
        if (!Ag) {
            Ag = malloc((size_t)M * N * sizeof(double));
            gen_global(ictxt, M, N, NB, seedA, Ag);
        }
        for (int i = 0; i < M; i++) {
            double s = 0.0;
            for (int j = 0; j < N; j++) s += fabs(Ag[i + (size_t)j * M]);
            if (s > anorm) anorm = s;
        }
        if (solve) {
            const double eps = ldexp(1.0, -53);
            sresid = 0.0;
            double *r = malloc((size_t)N * sizeof(double));
            for (int c = 0; c < NRHS; c++) {
                const double *x = Xg + (size_t)c * N;
                for (int i = 0; i < N; i++) r[i] = Bg[i + (size_t)c * N];
                for (int j = 0; j < N; j++) {
                    double xj = x[j];
                    const double *a = Ag + (size_t)j * M;
                    for (int i = 0; i < N; i++) r[i] -= a[i] * xj;
                }
                double rn = 0.0, xn = 0.0;
                for (int i = 0; i < N; i++) {
                    if (fabs(r[i]) > rn) rn = fabs(r[i]);
                    if (fabs(x[i]) > xn) xn = fabs(x[i]);
                }
                double s = rn / (xn * anorm * eps * N);
                if (s > sresid || s != s) sresid = s;
            }
            free(r);
        }
        FILE *f = fopen(out, "wb");
        int32_t hdr[10] = {0x4C55524E, M, N, NB, NRHS, P, Q, info_f, info_s,
                           (solve ? 1 : 0) | (want_lu ? 2 : 0)};
        double dh[4] = {t_getrf, t_getrs, sresid, anorm};
        fwrite(hdr, sizeof hdr, 1, f);
        fwrite(dh, sizeof dh, 1, f);
        fwrite(ipg, sizeof(int), MN, f);
        if (solve) fwrite(Xg, sizeof(double), (size_t)N * NRHS, f);
        if (want_lu) fwrite(LUg, sizeof(double), (size_t)M * N, f);
        fclose(f);
        printf("{\"M\": %d, \"N\": %d, \"NB\": %d, \"NRHS\": %d, \"P\": %d, \"Q\": %d, "
               "\"info_getrf\": %d, \"info_getrs\": %d, \"t_getrf\": %.6f, \"t_getrs\": %.6f, "
               "\"sresid\": %.6g}\n",
               M, N, NB, NRHS, P, Q, info_f, info_s, t_getrf, t_getrs, sresid);
    }
    free(A); free(B); free(ipiv); free(ipg); free(Ag); free(Bg); free(Xg); free(LUg);
    blacs_gridexit_(&ictxt);
    int cont = 0;
    blacs_exit_(&cont);
    return 0;
}
