/*
 * checks.c -- the reference tester's residual checks, restated
 * (TEST INFRASTRUCTURE ONLY, see oracle.h).
 *   oracle_sresid: TESTING/traditional/LIN/pdlaschk.f:185-187,274-299
 *       max over RHS j of ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf eps N)
 *   oracle_fresid: TESTING/traditional/LIN/pdlafchk.f:199-226 after the
 *       P*L*U rebuild of pdgetrrv.f:210-289
 *       ||A - P L U||_inf / (max(M,N) eps ||A||_inf)
 * eps = DLAMCH('E') = 2^-53 (SRC/pdlamch.f:68-80, round-to-nearest).
 */
#include "oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

extern void dgemm_(const char *, const char *, const int *, const int *,
                   const int *, const double *, const double *, const int *,
                   const double *, const int *, const double *, double *,
                   const int *, size_t, size_t);

static const double EPS = 1.1102230246251565e-16; /* 2^-53 */

double oracle_norm_inf(int m, int n, const double *a, int lda)
{
    double *rs = calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) rs[i] += fabs(a[i + (size_t)j * lda]);
    double r = 0.0;
    for (int i = 0; i < m; i++)
        if (rs[i] > r || rs[i] != rs[i]) r = rs[i];
    free(rs);
    return r;
}

double oracle_sresid(int n, int nrhs, const double *a, int lda, const double *x,
                     int ldx, const double *b, int ldb)
{
    double anorm = oracle_norm_inf(n, n, a, lda);
    double *r = malloc((size_t)n * nrhs * sizeof(double));
    for (int j = 0; j < nrhs; j++)
        memcpy(r + (size_t)j * n, b + (size_t)j * ldb, (size_t)n * sizeof(double));
    const double one = 1.0, mone = -1.0;
    dgemm_("N", "N", &n, &nrhs, &n, &mone, a, &lda, x, &ldx, &one, r, &n, 1, 1);
    double res = 0.0;
    for (int j = 0; j < nrhs; j++) {
        double rn = 0.0, xn = 0.0;
        for (int i = 0; i < n; i++) {
            double rv = fabs(r[i + (size_t)j * n]), xv = fabs(x[i + (size_t)j * ldx]);
            if (rv > rn || rv != rv) rn = rv;
            if (xv > xn) xn = xv;
        }
        double s = rn / (xn * anorm * EPS * n);
        if (s > res || s != s) res = s;
    }
    free(r);
    return res;
}

double oracle_fresid(int m, int n, const double *a0, int lda0, const double *lu,
                     int ldlu, const int *ipiv)
{
    int mn = m < n ? m : n;
    double *L = calloc((size_t)m * mn, sizeof(double));
    double *U = calloc((size_t)mn * n, sizeof(double));
    double *R = malloc((size_t)m * n * sizeof(double));
    for (int j = 0; j < mn; j++)
        for (int i = 0; i < m; i++) {
            double v = lu[i + (size_t)j * ldlu];
            L[i + (size_t)j * m] = i > j ? v : (i == j ? 1.0 : 0.0);
        }
    for (int j = 0; j < n; j++)
        for (int i = 0; i <= j && i < mn; i++) U[i + (size_t)j * mn] = lu[i + (size_t)j * ldlu];
    const double one = 1.0, zero = 0.0;
    dgemm_("N", "N", &m, &n, &mn, &one, L, &m, U, &mn, &zero, R, &m, 1, 1);
    for (int i = mn - 1; i >= 0; i--) {          /* apply P: backward swaps */
        int ip = ipiv[i] - 1;
        if (ip != i)
            for (int j = 0; j < n; j++) {
                double t = R[i + (size_t)j * m];
                R[i + (size_t)j * m] = R[ip + (size_t)j * m];
                R[ip + (size_t)j * m] = t;
            }
    }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) R[i + (size_t)j * m] -= a0[i + (size_t)j * lda0];
    double rn = oracle_norm_inf(m, n, R, m);
    double an = oracle_norm_inf(m, n, a0, lda0);
    free(L); free(U); free(R);
    int mx = m > n ? m : n;
    return rn / (mx * EPS * an);
}
