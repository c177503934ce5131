/*
 * lu_oracle.c -- CPU restatement of the reference's 1 x 1 BLACS-grid LU path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h); the product never calls it.
 *
 * On a 1 x 1 grid every PBLAS call of the path reduces to one local BLAS call
 * (PBLAS/SRC/pdamax_.c:421-425, pdswap_.c:815 -> PTOOLS/PB_CpswapNN.c:267,
 * pdscal_.c:246, pdger_.c:288, pdtrsm_.c:497 -> PTOOLS/PB_CptrsmAB.c,
 * pdgemm_.c:476 -> PTOOLS/PB_CpgemmAB.c:333-352), so this file issues exactly
 * that call sequence against the host BLAS.
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

extern int idamax_(const int *, const double *, const int *);
extern void dswap_(const int *, double *, const int *, double *, const int *);
extern void dscal_(const int *, const double *, double *, const int *);
extern void dger_(const int *, const int *, const double *, const double *,
                  const int *, const double *, const int *, double *,
                  const int *);
extern void dtrsm_(const char *, const char *, const char *, const char *,
                   const int *, const int *, const double *, const double *,
                   const int *, double *, const int *, size_t, size_t, size_t,
                   size_t);
extern void dgemm_(const char *, const char *, const int *, const int *,
                   const int *, const double *, const double *, const int *,
                   const double *, const int *, const double *, double *,
                   const int *, size_t, size_t);

enum { LLD_ = 8, NB_ = 5, MB_ = 4 };
/* PILAENV for double precision (PBLAS/SRC/pilaenv.f:75). */
static const int PB_KB = 32;

#define AT(a, lda, i, j) ((a) + (size_t)(i) + (size_t)(j) * (size_t)(lda))

/* Relative gap between the two largest |x| of a column: the near-tie
 * indicator used to trace pivot divergences. */
static double top2_gap(int n, const double *x)
{
    double p1 = 0.0, p2 = 0.0;
    for (int i = 0; i < n; i++) {
        double v = fabs(x[i]);
        if (v > p1) { p2 = p1; p1 = v; }
        else if (v > p2) p2 = v;
    }
    return p1 > 0.0 ? (p1 - p2) / p1 : 0.0;
}

/* PDGETF2, SRC/pdgetf2.f:206-237 (1 x 1 grid). a is the local array, ia/ja
 * 1-based global indices of the panel, ipiv indexed by global row. */
int oracle_pdgetf2(int m, int n, double *a, int ia, int ja, const int *desca,
                   int *ipiv, double *gap)
{
    const int lda = desca[LLD_], one = 1;
    const double mone = -1.0;
    int info = 0;
    if (m == 0 || n == 0) return 0;
    int mn = m < n ? m : n;
    for (int j = 0; j < mn; j++) {
        int gi = ia + j;                   /* global row I of column J */
        double *col = AT(a, lda, gi - 1, ja - 1 + j);
        int len = m - j;
        int idx = idamax_(&len, col, &one);
        double gmax = col[idx - 1];
        if (gap) gap[j] = top2_gap(len, col);
        /* PDAMAX returns INDX = IX when AMAX is zero (pdamax_.c:486). */
        ipiv[gi - 1] = (gmax == 0.0) ? gi : gi + idx - 1;
        if (gmax != 0.0) {
            dswap_(&n, AT(a, lda, gi - 1, ja - 1), &lda,
                   AT(a, lda, ipiv[gi - 1] - 1, ja - 1), &lda);
            if (j + 1 < m) {
                int ns = m - j - 1;
                double r = 1.0 / gmax;     /* PDSCAL(..., ONE / GMAX, ...) */
                dscal_(&ns, &r, col + 1, &one);
            }
        } else if (info == 0) {
            info = j + 1;
        }
        if (j + 1 < mn) {
            int mm = m - j - 1, nn = n - j - 1;
            dger_(&mm, &nn, &mone, col + 1, &one, AT(a, lda, gi - 1, ja + j),
                  &lda, AT(a, lda, gi, ja + j), &lda);
        }
    }
    return info;
}

/* PDLASWP('F'|'B', 'R', ...), SRC/pdlaswp.f:163-185 (1 x 1 grid). */
void oracle_pdlaswp(char direc, int n, double *a, int ia, int ja,
                    const int *desca, int k1, int k2, const int *ipiv)
{
    (void)ia;
    const int lda = desca[LLD_];
    if (n == 0) return;
    if (direc == 'F' || direc == 'f') {
        for (int i = k1; i <= k2; i++) {
            int ip = ipiv[i - 1];
            if (ip != i)
                dswap_(&n, AT(a, lda, i - 1, ja - 1), &lda,
                       AT(a, lda, ip - 1, ja - 1), &lda);
        }
    } else {
        for (int i = k2; i >= k1; i--) {
            int ip = ipiv[i - 1];
            if (ip != i)
                dswap_(&n, AT(a, lda, i - 1, ja - 1), &lda,
                       AT(a, lda, ip - 1, ja - 1), &lda);
        }
    }
}

/* PDTRSM('L', 'L', 'N', 'U') on a 1 x 1 grid: PB_CptrsmAB's forward loop in
 * blocks of PILAENV = 32 rows, a local dtrsm_ per block
 * (PTOOLS/PB_CptrsmAB0.c:85-99) and a dgemm_ for the rows below
 * (PTOOLS/PB_CptrsmAB.c:359-412). */
static void trsm_llnu(int m, int n, const double *l, int ldl, double *b, int ldb)
{
    const double one = 1.0, mone = -1.0;
    for (int k = 0; k < m; k += PB_KB) {
        int kbb = m - k < PB_KB ? m - k : PB_KB;
        dtrsm_("L", "L", "N", "U", &kbb, &n, &one, AT(l, ldl, k, k), &ldl,
               AT(b, ldb, k, 0), &ldb, 1, 1, 1, 1);
        int rest = m - k - kbb;
        if (rest > 0)
            dgemm_("N", "N", &rest, &n, &kbb, &mone, AT(l, ldl, k + kbb, k), &ldl,
                   AT(b, ldb, k, 0), &ldb, &one, AT(b, ldb, k + kbb, 0), &ldb, 1, 1);
    }
}

/* PDTRSM('L', 'U', 'N', 'N') on a 1 x 1 grid: PB_CptrsmAB's backward loop
 * (PTOOLS/PB_CptrsmAB.c:307-355). */
static void trsm_lunn(int m, int n, const double *u, int ldu, double *b, int ldb)
{
    const double one = 1.0, mone = -1.0;
    int kmax = ((m - 1) / PB_KB) * PB_KB;
    for (int k = kmax; k >= 0; k -= PB_KB) {
        int kbb = m - k < PB_KB ? m - k : PB_KB;
        dtrsm_("L", "U", "N", "N", &kbb, &n, &one, AT(u, ldu, k, k), &ldu,
               AT(b, ldb, k, 0), &ldb, 1, 1, 1, 1);
        if (k > 0)
            dgemm_("N", "N", &k, &n, &kbb, &mone, AT(u, ldu, 0, k), &ldu,
                   AT(b, ldb, k, 0), &ldb, &one, b, &ldb, 1, 1);
    }
}

/* PDGETRF, SRC/pdgetrf.f:217-295 (1 x 1 grid, IA/JA aligned to NB). */
int oracle_pdgetrf(int m, int n, double *a, int ia, int ja, const int *desca,
                   int *ipiv, double *gap)
{
    const int lda = desca[LLD_], nb = desca[NB_];
    const double one = 1.0, mone = -1.0;
    int info = 0;
    if (desca[2] == 1) { ipiv[0] = 1; return 0; }   /* pdgetrf.f:201-203 */
    if (m == 0 || n == 0) return 0;
    int mn = m < n ? m : n;
    int jn = ((ja + nb - 1) / nb) * nb;              /* ICEIL(JA,NB)*NB */
    if (jn > ja + mn - 1) jn = ja + mn - 1;
    int in = ((ia + nb - 1) / nb) * nb;
    if (in > ia + m - 1) in = ia + m - 1;
    int jb = jn - ja + 1;

    info = oracle_pdgetf2(m, jb, a, ia, ja, desca, ipiv, gap);
    if (jb + 1 <= n) {
        oracle_pdlaswp('F', n - jb, a, ia, jn + 1, desca, ia, in, ipiv);
        int nr = n - jb;
        trsm_llnu(jb, nr, AT(a, lda, ia - 1, ja - 1), lda,
                  AT(a, lda, ia - 1, jn), lda);
        if (jb + 1 <= m) {
            int mr = m - jb;
            dgemm_("N", "N", &mr, &nr, &jb, &mone, AT(a, lda, in, ja - 1), &lda,
                   AT(a, lda, ia - 1, jn), &lda, &one, AT(a, lda, in, jn), &lda, 1, 1);
        }
    }
    for (int j = jn + 1; j <= ja + mn - 1; j += nb) {
        int jbb = mn - j + ja < nb ? mn - j + ja : nb;
        int i = ia + j - ja;
        int iinfo = oracle_pdgetf2(m - j + ja, jbb, a, i, j, desca, ipiv,
                                   gap ? gap + (j - ja) : NULL);
        if (info == 0 && iinfo > 0) info = iinfo + j - ja;
        oracle_pdlaswp('F', j - ja, a, ia, ja, desca, i, i + jbb - 1, ipiv);
        if (j - ja + jbb + 1 <= n) {
            int nr = n - j - jbb + ja;
            oracle_pdlaswp('F', nr, a, ia, j + jbb, desca, i, i + jbb - 1, ipiv);
            trsm_llnu(jbb, nr, AT(a, lda, i - 1, j - 1), lda,
                      AT(a, lda, i - 1, j + jbb - 1), lda);
            if (j - ja + jbb + 1 <= m) {
                int mr = m - j - jbb + ja;
                dgemm_("N", "N", &mr, &nr, &jbb, &mone, AT(a, lda, i + jbb - 1, j - 1),
                       &lda, AT(a, lda, i - 1, j + jbb - 1), &lda, &one,
                       AT(a, lda, i + jbb - 1, j + jbb - 1), &lda, 1, 1);
            }
        }
    }
    return info;
}

/* PDGETRS('N'), SRC/pdgetrs.f:244-266: PDLAPIV forward on the rows of B
 * (SRC/pdlapv2.f:208-242, IP1 = IPIV - IA + IB), then L and U solves. */
int oracle_pdgetrs(char trans, int n, int nrhs, const double *a, int ia, int ja,
                   const int *desca, const int *ipiv, double *b, int ib, int jb,
                   const int *descb)
{
    const int lda = desca[LLD_], ldb = descb[LLD_];
    if (!(trans == 'N' || trans == 'n')) return -1;
    if (n == 0 || nrhs == 0) return 0;
    for (int k = ib; k <= ib + n - 1; k++) {
        int ip1 = ipiv[ia + (k - ib) - 1] - ia + ib;
        if (ip1 != k)
            dswap_(&nrhs, AT(b, ldb, k - 1, jb - 1), &ldb,
                   AT(b, ldb, ip1 - 1, jb - 1), &ldb);
    }
    trsm_llnu(n, nrhs, AT(a, lda, ia - 1, ja - 1), lda, AT(b, ldb, ib - 1, jb - 1), ldb);
    trsm_lunn(n, nrhs, AT(a, lda, ia - 1, ja - 1), lda, AT(b, ldb, ib - 1, jb - 1), ldb);
    return 0;
}

/* PDGESV, SRC/pdgesv.f:229-236. */
int oracle_pdgesv(int n, int nrhs, double *a, int ia, int ja, const int *desca,
                  int *ipiv, double *b, int ib, int jb, const int *descb)
{
    int info = oracle_pdgetrf(n, n, a, ia, ja, desca, ipiv, NULL);
    if (info == 0)
        info = oracle_pdgetrs('N', n, nrhs, a, ia, ja, desca, ipiv, b, ib, jb, descb);
    return info;
}
