// sla_aux.h -- norms, condition estimate and iterative refinement with the
// factors of PDGETRF (aux.hip), on device-resident data.
#pragma once

#include "sla_handle.h"

namespace sla {

// PDLANGE: 'M' max |a|, '1'/'O' one norm, 'I' infinity norm, 'F'/'E'
// Frobenius norm of the m x n matrix at dA.
double dlange_device(Handle &h, char norm, int m, int n, const double *dA, int lda);

// PDGECON: reciprocal condition number estimate in the one ('1'/'O') or
// infinity ('I') norm from the L\U factors at dLU and the norm of A.
double dgecon_device(Handle &h, char norm, int n, const double *dLU, int lda, double anorm);

// PDGERFS: refines the nrhs solutions at dX of op(A) X = B (op = transpose
// when trans) with the factors dAF / dipiv of A; writes the forward error
// bounds ferr[k] (left unchanged when x_k is zero) and the componentwise
// backward errors berr[k] (host arrays).
void dgerfs_device(Handle &h, bool trans, int n, int nrhs, const double *dA, int lda,
                   const double *dAF, int ldaf, const int *dipiv, const double *dB, int ldb,
                   double *dX, int ldx, double *ferr, double *berr);

// PDGEEQU: row scales dR[m], column scales dC[n] (device), the ratios and
// max |a|; returns INFO (i: row i exactly zero; m + j: column j).
int dgeequ_device(Handle &h, int m, int n, const double *dA, int lda, double *dR, double *dC,
                  double *rowcnd, double *colcnd, double *amax);
// PDLAQGE: scales dA in place as the ratios ask; returns EQUED.
char dlaqge_device(Handle &h, int m, int n, double *dA, int lda, const double *dR, const double *dC,
                   double rowcnd, double colcnd, double amax);
// PDGESVX with FACT 'N', 'E' or 'F' (equed in/out); returns INFO (n + 1:
// RCOND below eps, X not computed).
int dgesvx_device(Handle &h, char fact, bool trans, int n, int nrhs, double *dA, int lda, double *dAF,
                  int ldaf, int *dipiv, char *equed, double *dR, double *dC, double *dB, int ldb,
                  double *dX, int ldx, double *rcond, double *ferr, double *berr);

}  // namespace sla
