// gemm_small.hip -- the thin GEMMs of the recursive panel and TRSM:
// C[m x n] -= A[m x k] * B[k x n] when two of the three sizes are small.
//
//   tall-narrow  (n, k <= 64, m large): the GEMM between the halves of the
//                recursive panel, C = A22 (m x n1) -= L21 (m x n1) U12 (n1 x n1).
//                A thread computes 16 outputs of one row of C (a 2-D grid
//                over 256-row slabs and 16-column groups); the 16 columns of
//                B sit in LDS and are read as broadcasts; A and C stream with
//                coalesced accesses.
//   short-wide   (m, k <= 64, n large): the updates inside the recursive TRSM
//                of a block row, C = A12 rows below (m x n) -= L (m x k) X (k x n).
//                A thread computes 16 outputs of one column of C (2-D grid
//                over 256-column slabs and 16-row groups); A's 16 rows sit in
//                LDS.
// Both are bound by HBM traffic (each element of the large operands moves
// once); the MFMA kernels of gemm_f64.hip / gemm_lowp.hip need 128-wide
// tiles in both dimensions and would run these shapes mostly empty.
#include "sla_common.h"
#include "sla_kernels.h"

namespace sla {

namespace {

constexpr int SM_THREADS = 256;
constexpr int SM_SUB = 16;  // outputs per thread along the small dimension

// Grid: x over 256-row slabs of C, y over 16-column groups of C.
template <typename T>
__global__ __launch_bounds__(SM_THREADS) void gemm_tallnarrow_kernel(
    int m, int n, int k, const T *__restrict__ A, int lda, const T *__restrict__ B, int ldb,
    T *__restrict__ C, int ldc)
{
    __shared__ T Bs[SMALL_KMAX * SM_SUB];  // row kk holds B(kk, j0..j0+15)
    const int j0 = blockIdx.y * SM_SUB;
    for (int idx = threadIdx.x; idx < k * SM_SUB; idx += SM_THREADS) {
        const int kk = idx / SM_SUB, j = idx % SM_SUB;
        Bs[idx] = (j0 + j < n) ? B[kk + (size_t)(j0 + j) * ldb] : T(0);
    }
    __syncthreads();
    const int row = blockIdx.x * SM_THREADS + threadIdx.x;
    if (row >= m) return;
    T acc[SM_SUB];
#pragma unroll
    for (int j = 0; j < SM_SUB; j++) acc[j] = T(0);
    const T *Ar = A + row;
#pragma unroll 4
    for (int kk = 0; kk < k; kk++) {
        const T a = Ar[(size_t)kk * lda];
        const T *b = Bs + kk * SM_SUB;
#pragma unroll
        for (int j = 0; j < SM_SUB; j++) acc[j] = fma(a, b[j], acc[j]);
    }
    T *Cr = C + row + (size_t)j0 * ldc;
#pragma unroll
    for (int j = 0; j < SM_SUB; j++)
        if (j0 + j < n) Cr[(size_t)j * ldc] -= acc[j];
}

// Grid: x over 256-column slabs of C, y over 16-row groups of C.
template <typename T>
__global__ __launch_bounds__(SM_THREADS) void gemm_shortwide_kernel(
    int m, int n, int k, const T *__restrict__ A, int lda, const T *__restrict__ B, int ldb,
    T *__restrict__ C, int ldc)
{
    __shared__ T As[SMALL_KMAX * SM_SUB];  // row kk holds A(i0..i0+15, kk)
    const int i0 = blockIdx.y * SM_SUB;
    for (int idx = threadIdx.x; idx < k * SM_SUB; idx += SM_THREADS) {
        const int kk = idx / SM_SUB, i = idx % SM_SUB;
        As[idx] = (i0 + i < m) ? A[i0 + i + (size_t)kk * lda] : T(0);
    }
    __syncthreads();
    const int col = blockIdx.x * SM_THREADS + threadIdx.x;
    if (col >= n) return;
    T acc[SM_SUB];
#pragma unroll
    for (int i = 0; i < SM_SUB; i++) acc[i] = T(0);
    const T *Bc = B + (size_t)col * ldb;
#pragma unroll 4
    for (int kk = 0; kk < k; kk++) {
        const T b = Bc[kk];
        const T *a = As + kk * SM_SUB;
#pragma unroll
        for (int i = 0; i < SM_SUB; i++) acc[i] = fma(a[i], b, acc[i]);
    }
    T *Cc = C + (size_t)col * ldc + i0;
#pragma unroll
    for (int i = 0; i < SM_SUB; i++)
        if (i0 + i < m) Cc[i] -= acc[i];
}

}  // namespace

template <typename T>
bool gemm_small(int m, int n, int k, const T *A, int lda, const T *B, int ldb, T *C, int ldc,
                hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return true;
    if (k > SMALL_KMAX) return false;
    if (n <= SMALL_NMAX && m >= 4 * n) {
        hipLaunchKernelGGL(gemm_tallnarrow_kernel<T>,
                           dim3(ceil_div(m, SM_THREADS), ceil_div(n, SM_SUB)), dim3(SM_THREADS),
                           0, s, m, n, k, A, lda, B, ldb, C, ldc);
        SLA_LAUNCH_CHECK();
        return true;
    }
    if (m <= SMALL_NMAX) {
        hipLaunchKernelGGL(gemm_shortwide_kernel<T>,
                           dim3(ceil_div(n, SM_THREADS), ceil_div(m, SM_SUB)), dim3(SM_THREADS),
                           0, s, m, n, k, A, lda, B, ldb, C, ldc);
        SLA_LAUNCH_CHECK();
        return true;
    }
    return false;
}

template bool gemm_small<double>(int, int, int, const double *, int, const double *, int, double *,
                                 int, hipStream_t);
template bool gemm_small<float>(int, int, int, const float *, int, const float *, int, float *, int,
                                hipStream_t);

}  // namespace sla
