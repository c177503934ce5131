/*
 * aux_oracle.c -- CPU restatement of the routines beside the 1 x 1 LU solve:
 * PDLANGE, PDLACON + PDGECON, PDGERFS and PDGETRS with TRANS = 'T'.
 * TEST INFRASTRUCTURE ONLY (see oracle.h); the product never calls it.
 *
 * On a 1 x 1 grid each PBLAS call below is one local call: PDASUM -> dasum_,
 * PDAMAX -> idamax_, PDGEMV -> dgemv_, PDTRSV / PDTRSM -> dtrsv_ / dtrsm_,
 * PDAGEMV -> the loops of PBLAS/SRC/PTZBLAS/dagemv.f, restated here.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

extern double dasum_(const int *, const double *, const int *);
extern int idamax_(const int *, const double *, const int *);
extern void dswap_(const int *, double *, const int *, double *, const int *);
extern void dlassq_(const int *, const double *, const int *, double *, double *);
extern void dtrsv_(const char *, const char *, const char *, const int *, const double *,
                   const int *, double *, const int *, size_t, size_t, size_t);
extern void dtrsm_(const char *, const char *, const char *, const char *, const int *,
                   const int *, const double *, const double *, const int *, double *,
                   const int *, size_t, size_t, size_t, size_t);
extern void dgemv_(const char *, const int *, const int *, const double *, const double *,
                   const int *, const double *, const int *, const double *, double *,
                   const int *, size_t);

enum { MB_ = 4, NB_ = 5, LLD_ = 8 };
#define AT(a, lda, i, j) ((a) + (size_t)(i) + (size_t)(j) * (size_t)(lda))

static int up(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }

/* DCOMBSSQ, TOOLS/pdtreecomb.f:260-280. */
static void combssq(double *v1, const double *v2)
{
    if (v1[0] >= v2[0]) {
        if (v1[0] != 0.0) v1[1] = v1[1] + (v2[0] / v1[0]) * (v2[0] / v1[0]) * v2[1];
    } else {
        v1[1] = v2[1] + (v1[0] / v2[0]) * (v1[0] / v2[0]) * v1[1];
        v1[0] = v2[0];
    }
}

/* PDLANGE, SRC/pdlange.f:196-330 (the 1 x 1 local loops). */
double oracle_pdlange(char norm, int m, int n, const double *a, int ia, int ja,
                      const int *desca)
{
    const int lda = desca[LLD_];
    const double *s = AT(a, lda, ia - 1, ja - 1);
    const int c = up(norm);
    if ((m < n ? m : n) == 0) return 0.0;
    if (c == 'M') {
        double v = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) v = fmax(v, fabs(*AT(s, lda, i, j)));
        return v;
    }
    if (c == 'O' || c == '1') {
        double *w = malloc((size_t)n * sizeof(double));
        for (int j = 0; j < n; j++) {
            double sum = 0.0;
            for (int i = 0; i < m; i++) sum = sum + fabs(*AT(s, lda, i, j));
            w[j] = sum;
        }
        const int one = 1;
        double v = w[idamax_(&n, w, &one) - 1];
        free(w);
        return v;
    }
    if (c == 'I') {
        double *w = malloc((size_t)m * sizeof(double));
        for (int i = 0; i < m; i++) {
            double sum = 0.0;
            for (int j = 0; j < n; j++) sum = sum + fabs(*AT(s, lda, i, j));
            w[i] = sum;
        }
        const int one = 1;
        double v = w[idamax_(&m, w, &one) - 1];
        free(w);
        return v;
    }
    if (c == 'F' || c == 'E') {
        double ssq[2] = {0.0, 1.0};
        const int one = 1;
        for (int j = 0; j < n; j++) {
            double col[2] = {0.0, 1.0};
            dlassq_(&m, AT(s, lda, 0, j), &one, &col[0], &col[1]);
            combssq(ssq, col);
        }
        return ssq[0] * sqrt(ssq[1]);
    }
    return 0.0;
}

/* PDLACON, SRC/pdlacon.f (jump labels 20 .. 140, ITMAX = 5), driven by a
 * callback in place of the reverse communication: apply(kase, x, ctx).
 * pdlacon.f:188-189 set EST = 0 and ESTWORK(1) = EST on every entry, so
 * across the returns to the caller the running estimate does not survive:
 * ESTOLD at label 70 is 0, and label 140 compares TEMP with 0.  Both are
 * restated as such (est_entry below is the value ESTWORK(1) holds after
 * that reset). */
typedef void (*apply_fn)(int kase, double *x, void *ctx);

static double lacon(int n, apply_fn apply, void *ctx)
{
    const int one = 1, itmax = 5;
    double *x = malloc((size_t)n * sizeof(double)), *v = malloc((size_t)n * sizeof(double));
    int *isgn = malloc((size_t)n * sizeof(int));
    double est;
    for (int i = 0; i < n; i++) x[i] = 1.0 / (double)n;
    apply(1, x, ctx);
    if (n == 1) {
        v[0] = x[0];
        est = fabs(v[0]);
        goto done;
    }
    est = dasum_(&n, x, &one);
    for (int i = 0; i < n; i++) {
        x[i] = copysign(1.0, x[i]);
        isgn[i] = (int)lrint(x[i]);
    }
    apply(2, x, ctx);
    int j = idamax_(&n, x, &one) - 1;
    int iter = 2;
    const double est_entry = 0.0;
    for (;;) {
        for (int i = 0; i < n; i++) x[i] = 0.0;
        x[j] = 1.0;
        apply(1, x, ctx);
        memcpy(v, x, (size_t)n * sizeof(double));
        const double estold = est_entry;  /* ESTOLD = ESTWORK(1) after the reset */
        est = dasum_(&n, v, &one);
        int changed = 0;
        for (int i = 0; i < n; i++)
            if ((int)lrint(copysign(1.0, x[i])) != isgn[i]) {
                changed = 1;
                break;
            }
        if (!changed || est <= estold) break;
        for (int i = 0; i < n; i++) {
            x[i] = copysign(1.0, x[i]);
            isgn[i] = (int)lrint(x[i]);
        }
        apply(2, x, ctx);
        const int jlast = j;
        j = idamax_(&n, x, &one) - 1;
        if (x[jlast] != fabs(x[j]) && iter < itmax) {
            iter++;
            continue;
        }
        break;
    }
    for (int i = 0; i < n; i++) {
        const int k = i + 1;
        const double altsgn = (k % 2 == 0) ? -1.0 : 1.0;
        x[i] = altsgn * (1.0 + (double)(k - 1) / (double)(n - 1));
    }
    apply(1, x, ctx);
    {
        double temp = dasum_(&n, x, &one);
        temp = 2.0 * (temp / (double)(3 * n));
        est = est_entry;  /* label 140 runs after an entry: ESTWORK(1) = 0 */
        if (temp > est) {
            memcpy(v, x, (size_t)n * sizeof(double));
            est = temp;
        }
    }
done:
    free(x);
    free(v);
    free(isgn);
    return est;
}

struct gecon_ctx {
    int n, lda, kase1;
    const double *lu;
};

static void gecon_apply(int kase, double *x, void *vctx)
{
    const struct gecon_ctx *c = vctx;
    const int one = 1;
    if (kase == c->kase1) {
        dtrsv_("L", "N", "U", &c->n, c->lu, &c->lda, x, &one, 1, 1, 1);
        dtrsv_("U", "N", "N", &c->n, c->lu, &c->lda, x, &one, 1, 1, 1);
    } else {
        dtrsv_("U", "T", "N", &c->n, c->lu, &c->lda, x, &one, 1, 1, 1);
        dtrsv_("L", "T", "U", &c->n, c->lu, &c->lda, x, &one, 1, 1, 1);
    }
}

/* PDGECON, SRC/pdgecon.f:287-403 (PDLATRS = PDTRSV with SCALE = 1,
 * SRC/pdlatrs.f:65-67). */
double oracle_pdgecon(char norm, int n, const double *a, int ia, int ja, const int *desca,
                      double anorm)
{
    if (n == 0) return 1.0;
    if (anorm == 0.0) return 0.0;
    if (n == 1) return 1.0;
    const int c = up(norm);
    struct gecon_ctx ctx = {n, desca[LLD_], (c == '1' || c == 'O') ? 1 : 2,
                            AT(a, desca[LLD_], ia - 1, ja - 1)};
    const double ainvnm = lacon(n, gecon_apply, &ctx);
    return ainvnm != 0.0 ? (1.0 / ainvnm) / anorm : 0.0;
}

/* PDGETRS on one right-hand side vector at x (contiguous): TRANS = 'N'
 * (pdgetrs.f:255-266) or 'T' (:273-284: U^T, L^T, then PDLAPIV backward). */
static void getrs_vec(int trans, int n, const double *lu, int lda, const int *piv, double *x)
{
    const int one = 1;
    if (!trans) {
        for (int k = 0; k < n; k++)
            if (piv[k] - 1 != k) dswap_(&one, x + k, &one, x + piv[k] - 1, &one);
        dtrsv_("L", "N", "U", &n, lu, &lda, x, &one, 1, 1, 1);
        dtrsv_("U", "N", "N", &n, lu, &lda, x, &one, 1, 1, 1);
    } else {
        dtrsv_("U", "T", "N", &n, lu, &lda, x, &one, 1, 1, 1);
        dtrsv_("L", "T", "U", &n, lu, &lda, x, &one, 1, 1, 1);
        for (int k = n - 1; k >= 0; k--)
            if (piv[k] - 1 != k) dswap_(&one, x + k, &one, x + piv[k] - 1, &one);
    }
}

int oracle_pdgetrs_t(int n, int nrhs, const double *a, int ia, int ja, const int *desca,
                     const int *ipiv, double *b, int ib, int jb, const int *descb)
{
    const int lda = desca[LLD_], ldb = descb[LLD_];
    if (n == 0 || nrhs == 0) return 0;
    const double one = 1.0;
    double *bs = AT(b, ldb, ib - 1, jb - 1);
    const double *as = AT(a, lda, ia - 1, ja - 1);
    dtrsm_("L", "U", "T", "N", &n, &nrhs, &one, as, &lda, bs, &ldb, 1, 1, 1, 1);
    dtrsm_("L", "L", "T", "U", &n, &nrhs, &one, as, &lda, bs, &ldb, 1, 1, 1, 1);
    for (int k = ib + n - 1; k >= ib; k--) {
        const int ip1 = ipiv[ia + (k - ib) - 1] - ia + ib;
        if (ip1 != k) dswap_(&nrhs, AT(b, ldb, k - 1, jb - 1), &ldb, AT(b, ldb, ip1 - 1, jb - 1), &ldb);
    }
    return 0;
}

struct rfs_ctx {
    int n, lda, trans;
    const double *lu;
    const int *piv;
    const double *w;
};

static void rfs_apply(int kase, double *x, void *vctx)
{
    const struct rfs_ctx *c = vctx;
    if (kase == 1) {
        getrs_vec(!c->trans, c->n, c->lu, c->lda, c->piv, x);
        for (int i = 0; i < c->n; i++) x[i] = c->w[i] * x[i];
    } else {
        for (int i = 0; i < c->n; i++) x[i] = c->w[i] * x[i];
        getrs_vec(c->trans, c->n, c->lu, c->lda, c->piv, x);
    }
}

/* PDGERFS, SRC/pdgerfs.f:440-675, with IA = JA = IAF = JAF = IB = JB = IX =
 * JX = 1: A, AF (lda), B (ldb), X (ldx), IPIV 1-based. */
int oracle_pdgerfs(char trans, int n, int nrhs, const double *a, int lda, const double *af,
                   int ldaf, const int *ipiv, const double *b, int ldb, double *x, int ldx,
                   double *ferr, double *berr)
{
    const int tr = up(trans) != 'N';
    const int itmax = 5, one = 1;
    if (n <= 1 || nrhs == 0) {
        for (int k = 0; k < nrhs; k++) ferr[k] = berr[k] = 0.0;
        return 0;
    }
    const int nz = n + 1;
    const double eps = ldexp(1.0, -53), safmin = DBL_MIN;
    const double safe1 = nz * safmin, safe2 = safe1 / eps;
    const double mone = -1.0, pone = 1.0;
    double *r = malloc((size_t)n * sizeof(double)), *w = malloc((size_t)n * sizeof(double));
    const char *tc = tr ? "T" : "N";
    for (int k = 0; k < nrhs; k++) {
        const double *bk = b + (size_t)k * ldb;
        double *xk = x + (size_t)k * ldx;
        int count = 1;
        double lstres = 3.0, s;
        for (;;) {
            memcpy(r, bk, (size_t)n * sizeof(double));
            dgemv_(tc, &n, &n, &mone, a, &lda, xk, &one, &pone, r, &one, 1);
            for (int i = 0; i < n; i++) w[i] = fabs(bk[i]);
            /* PDAGEMV(TRANS, 1, A, X, 1, W) = dagemv.f loops with BETA = 1 */
            if (!tr) {
                for (int j = 0; j < n; j++) {
                    const double absx = fabs(xk[j]);
                    if (absx != 0.0)
                        for (int i = 0; i < n; i++) w[i] = w[i] + absx * fabs(a[i + (size_t)j * lda]);
                }
            } else {
                for (int j = 0; j < n; j++) {
                    double temp = 0.0;
                    for (int i = 0; i < n; i++) temp = temp + fabs(a[i + (size_t)j * lda] * xk[i]);
                    w[j] = w[j] + temp;
                }
            }
            s = 0.0;
            for (int i = 0; i < n; i++) {
                const double t = w[i] > safe2 ? fabs(r[i]) / w[i] : (fabs(r[i]) + safe1) / (w[i] + safe1);
                s = fmax(s, t);
            }
            berr[k] = s;
            if (s > eps && 2.0 * s <= lstres && count <= itmax) {
                getrs_vec(tr, n, af, ldaf, ipiv, r);
                for (int i = 0; i < n; i++) xk[i] = xk[i] + r[i];
                lstres = s;
                count++;
                continue;
            }
            break;
        }
        for (int i = 0; i < n; i++)
            w[i] = w[i] > safe2 ? fabs(r[i]) + nz * eps * w[i] : fabs(r[i]) + nz * eps * w[i] + safe1;
        struct rfs_ctx ctx = {n, ldaf, tr, af, ipiv, w};
        const double est = lacon(n, rfs_apply, &ctx);
        double xmax = 0.0;
        for (int i = 0; i < n; i++) xmax = fmax(xmax, fabs(xk[i]));
        if (xmax != 0.0) ferr[k] = est / xmax;
    }
    free(r);
    free(w);
    return 0;
}

/* PDGEEQU, SRC/pdgeequ.f (1 x 1, IA = JA = 1): row scales R, column scales
 * C, ROWCND, COLCND, AMAX; INFO = i (row i exactly zero) or M + j. */
int oracle_pdgeequ(int m, int n, const double *a, int lda, double *r, double *c, double *rowcnd,
                   double *colcnd, double *amax)
{
    const double smlnum = DBL_MIN, bignum = 1.0 / smlnum;
    if (m == 0 || n == 0) {
        *rowcnd = 1.0;
        *colcnd = 1.0;
        *amax = 0.0;
        return 0;
    }
    for (int i = 0; i < m; i++) r[i] = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) r[i] = fmax(r[i], fabs(a[i + (size_t)j * lda]));
    double rcmin = bignum, rcmax = 0.0;
    for (int i = 0; i < m; i++) {
        rcmax = fmax(rcmax, r[i]);
        rcmin = fmin(rcmin, r[i]);
    }
    *amax = rcmax;
    if (rcmin == 0.0) {
        for (int i = 0; i < m; i++)
            if (r[i] == 0.0) return i + 1;
    }
    for (int i = 0; i < m; i++) r[i] = 1.0 / fmin(fmax(r[i], smlnum), bignum);
    *rowcnd = fmax(rcmin, smlnum) / fmin(rcmax, bignum);
    for (int j = 0; j < n; j++) {
        c[j] = 0.0;
        for (int i = 0; i < m; i++) c[j] = fmax(c[j], fabs(a[i + (size_t)j * lda]) * r[i]);
    }
    rcmin = bignum;
    rcmax = 0.0;
    for (int j = 0; j < n; j++) {
        rcmin = fmin(rcmin, c[j]);
        rcmax = fmax(rcmax, c[j]);
    }
    if (rcmin == 0.0) {
        for (int j = 0; j < n; j++)
            if (c[j] == 0.0) return m + j + 1;
    }
    for (int j = 0; j < n; j++) c[j] = 1.0 / fmin(fmax(c[j], smlnum), bignum);
    *colcnd = fmax(rcmin, smlnum) / fmin(rcmax, bignum);
    return 0;
}

/* PDLAQGE, SRC/pdlaqge.f:205-262 (THRESH = 0.1, SMALL = SAFMIN / PRECISION
 * with PRECISION = EPS * BASE = 2^-52); returns EQUED. */
char oracle_pdlaqge(int m, int n, double *a, int lda, const double *r, const double *c,
                    double rowcnd, double colcnd, double amax)
{
    const double thresh = 0.1, small = DBL_MIN / ldexp(1.0, -52), large = 1.0 / small;
    if (m <= 0 || n <= 0) return 'N';
    if (rowcnd >= thresh && amax >= small && amax <= large) {
        if (colcnd >= thresh) return 'N';
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) a[i + (size_t)j * lda] = c[j] * a[i + (size_t)j * lda];
        return 'C';
    }
    if (colcnd >= thresh) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) a[i + (size_t)j * lda] = r[i] * a[i + (size_t)j * lda];
        return 'R';
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) a[i + (size_t)j * lda] = c[j] * r[i] * a[i + (size_t)j * lda];
    return 'B';
}

/* PDGESVX, SRC/pdgesvx.f:624-737 for FACT = 'E' or 'N' (IA = JA = IB = JB =
 * IX = JX = 1, NB the block size of PDGETRF): equilibrate, factor, RCOND by
 * PDGECON, INFO = N + 1 when RCOND < EPS, solve, refine, undo the scaling
 * of X and FERR.  A, B are overwritten as the reference does. */
int oracle_pdgesvx(char fact, char trans, int n, int nrhs, int nb, double *a, double *af,
                   int *ipiv, char *equed, double *r, double *c, double *b, double *x,
                   double *rcond, double *ferr, double *berr)
{
    const int notran = up(trans) == 'N';
    int rowequ = 0, colequ = 0;
    double rowcnd = 1.0, colcnd = 1.0, amax = 0.0;
    *equed = 'N';
    if (up(fact) == 'E') {
        if (oracle_pdgeequ(n, n, a, n, r, c, &rowcnd, &colcnd, &amax) == 0) {
            *equed = oracle_pdlaqge(n, n, a, n, r, c, rowcnd, colcnd, amax);
            rowequ = *equed == 'R' || *equed == 'B';
            colequ = *equed == 'C' || *equed == 'B';
        }
    }
    if (notran) {
        if (rowequ)
            for (int j = 0; j < nrhs; j++)
                for (int i = 0; i < n; i++) b[i + (size_t)j * n] = r[i] * b[i + (size_t)j * n];
    } else if (colequ) {
        for (int j = 0; j < nrhs; j++)
            for (int i = 0; i < n; i++) b[i + (size_t)j * n] = c[i] * b[i + (size_t)j * n];
    }
    int desca[9] = {1, 0, n, n, nb, nb, 0, 0, n > 1 ? n : 1};
    memcpy(af, a, (size_t)n * n * sizeof(double));
    int info = oracle_pdgetrf(n, n, af, 1, 1, desca, ipiv, NULL);
    if (info != 0) {
        if (info > 0) *rcond = 0.0;
        return info;
    }
    const char norm = notran ? '1' : 'I';
    const double anorm = oracle_pdlange(norm, n, n, a, 1, 1, desca);
    *rcond = oracle_pdgecon(norm, n, af, 1, 1, desca, anorm);
    if (*rcond < ldexp(1.0, -53)) return n + 1;
    memcpy(x, b, (size_t)n * nrhs * sizeof(double));
    int descb[9] = {1, 0, n, nrhs, nb, nb, 0, 0, n > 1 ? n : 1};
    if (notran) oracle_pdgetrs('N', n, nrhs, af, 1, 1, desca, ipiv, x, 1, 1, descb);
    else oracle_pdgetrs_t(n, nrhs, af, 1, 1, desca, ipiv, x, 1, 1, descb);
    oracle_pdgerfs(trans, n, nrhs, a, n, af, n, ipiv, b, n, x, n, ferr, berr);
    if (notran) {
        if (colequ) {
            for (int j = 0; j < nrhs; j++)
                for (int i = 0; i < n; i++) x[i + (size_t)j * n] = c[i] * x[i + (size_t)j * n];
            for (int j = 0; j < nrhs; j++) ferr[j] = ferr[j] / colcnd;
        }
    } else if (rowequ) {
        for (int j = 0; j < nrhs; j++)
            for (int i = 0; i < n; i++) x[i + (size_t)j * n] = r[i] * x[i + (size_t)j * n];
        for (int j = 0; j < nrhs; j++) ferr[j] = ferr[j] / rowcnd;
    }
    return 0;
}
