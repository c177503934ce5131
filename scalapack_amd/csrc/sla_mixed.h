// sla_mixed.h -- the mixed-precision solve (low-precision LU + refinement to
// FP64 backward error) on device memory.
#pragma once

#include "sla_handle.h"

namespace sla {

// X = A^{-1} B for n x n A (device, FP64, column-major) and nrhs <= 16
// right-hand sides; see mixed.hip for the contract.  Returns INFO (> 0: the
// FP64 fallback found an exactly zero pivot).
int dxgesv_device(Handle &h, int prec, int refine, int n, int nrhs, double *dA, int lda,
                  int *dipiv, const double *dB, int ldb, double *dX, int ldx, MixedInfo &info);

// R = B - A X in FP64 (reproducible two-stage GEMV), rnorm[j] = max_i |R(i,j)|.
void dresidual_device(Handle &h, int n, int nrhs, const double *dA, int lda, const double *dX,
                      int ldx, const double *dB, int ldb, double *dR, int ldr, double *rnorm);
// B = A^{-1} B with FP32 factors of sgetrf_device.
void sgetrs_device(Handle &h, int n, int nrhs, const float *dA, int lda, const int *dipiv,
                   float *dB, int ldb);

}  // namespace sla
