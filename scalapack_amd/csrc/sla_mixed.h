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

}  // namespace sla
