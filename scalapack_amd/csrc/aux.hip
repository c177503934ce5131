// aux.hip -- the routines beside the solve, on device-resident data:
//
//   dlange_device  PDLANGE (SRC/pdlange.f:196-330): max |a|, one norm, infinity
//                  norm, Frobenius norm.  The one and infinity norms keep the
//                  reference's summation order (each column / row summed from
//                  its first element to its last, pdlange.f:228-236 and
//                  :268-276), so they agree with it bit for bit; the
//                  Frobenius norm scales by max |a| instead of DLASSQ's
//                  running scale and agrees to rounding.
//   lacon          PDLACON (SRC/pdlacon.f): Hager's 1-norm estimator with
//                  Higham's alternating-sign vector, written as a straight
//                  loop around a callback where the reference returns to its
//                  caller with KASE = 1 or 2.
//   dgecon_device  PDGECON (SRC/pdgecon.f:287-403): estimate of 1 / (||A||
//                  ||A^{-1}||) from PDGETRF's L and U with PDLATRS's solves,
//                  which on a 1 x 1 grid are plain triangular solves
//                  (SRC/pdlatrs.f:65-67, SCALE = 1).
//   dgerfs_device  PDGERFS (SRC/pdgerfs.f:473-675): per right-hand side the
//                  FP64 residual and |b| + |op(A)||x| in one pass over A, the
//                  componentwise backward error S, refinement steps while
//                  S > eps, 2 S <= last S and at most ITMAX = 5 steps, then
//                  the forward error bound by the estimator.
//
// Vectors, partial sums and reduction results live in one reusable device
// buffer of the handle; reductions run in a fixed order, so every result is
// reproducible from run to run.
#include "sla_common.h"
#include "sla_handle.h"
#include "sla_kernels.h"
#include "sla_aux.h"
#include "sla_batched.h"
#include "sla_mlp.h"

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace sla {

namespace {

constexpr double EPS64 = 1.1102230246251565e-16;  // 2^-53, PDLAMCH('Epsilon')
constexpr int RED_BLOCKS = 256;                    // stage-1 blocks of a vector reduction
constexpr int RF_ROWS = 256, RF_KC = 512;           // PDGERFS 'N': rows per block, column slice
constexpr int RF_TCOLS = 64, RF_TROWS = 1024;       // PDGERFS 'T': columns per block, row slice

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
// Non-negative doubles order like their bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *p, double v)
{
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// ---- PDLANGE ---------------------------------------------------------------
// max |A(i, j)|: x over rows, y over columns (strided).
__global__ __launch_bounds__(256) void lange_max_kernel(int m, int n, const double *__restrict__ A,
                                                        int lda, unsigned long long *out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < m)
        for (int j = blockIdx.y; j < n; j += gridDim.y) v = fmax(v, fabs(A[i + (size_t)j * lda]));
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, v);
}

// Column sums of |A| in the reference's order: one wave per 64 columns; each
// 64 x 64 tile is read down its columns (coalesced) into LDS, then lane c adds
// the tile's rows of column c one after another.
__global__ __launch_bounds__(64) void lange_colsum_kernel(int m, int n, const double *__restrict__ A,
                                                          int lda, double *__restrict__ colsum)
{
    __shared__ double T[64][65];
    const int lane = threadIdx.x, c0 = blockIdx.x * 64;
    const int cols = min(64, n - c0);
    double sum = 0.0;
    for (int r0 = 0; r0 < m; r0 += 64) {
        const int rows = min(64, m - r0);
        for (int c = 0; c < cols; c++)
            T[c][lane] = lane < rows ? fabs(A[(size_t)(r0 + lane) + (size_t)(c0 + c) * lda]) : 0.0;
        __syncthreads();
        if (lane < cols)
            for (int k = 0; k < rows; k++) sum += T[lane][k];
        __syncthreads();
    }
    if (lane < cols) colsum[c0 + lane] = sum;
}

// Row sums of |A| in the reference's order: thread i adds row i from column 0
// to column n-1 (consecutive threads read consecutive rows of each column).
__global__ __launch_bounds__(64) void lange_rowsum_kernel(int m, int n, const double *__restrict__ A,
                                                          int lda, double *__restrict__ rowsum)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= m) return;
    const double *a = A + i;
    double sum = 0.0;
#pragma unroll 8
    for (int j = 0; j < n; j++) sum += fabs(a[(size_t)j * lda]);
    rowsum[i] = sum;
}

// sum (A(i, j) / scale)^2 per block into part[block] (x over rows, y over
// column groups), scale = max |A| read from device memory.
__global__ __launch_bounds__(256) void lange_ssq_kernel(int m, int n, const double *__restrict__ A,
                                                        int lda, const unsigned long long *scale_bits,
                                                        double *__restrict__ part)
{
    __shared__ double ws[4];
    const double scale = __longlong_as_double((long long)*scale_bits);
    const double rs = scale > 0.0 ? 1.0 / scale : 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < m)
        for (int j = blockIdx.y; j < n; j += gridDim.y) {
            const double t = A[i + (size_t)j * lda] * rs;
            v = fma(t, t, v);
        }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// ---- vector reductions (two stages, fixed order) ---------------------------
enum RedOp { RED_SUM_ABS = 0, RED_MAX_ABS = 1, RED_SUM = 2, RED_MIN_ABS = 3 };

template <int OP>
__device__ __forceinline__ double red_combine(double a, double b)
{
    return OP == RED_MAX_ABS ? fmax(a, b) : OP == RED_MIN_ABS ? fmin(a, b) : a + b;
}
template <int OP>
__device__ __forceinline__ double red_init()
{
    return OP == RED_MIN_ABS ? INFINITY : 0.0;
}
template <int OP>
__device__ __forceinline__ double red_load(double v)
{
    return OP == RED_SUM ? v : fabs(v);
}

// Stage 1: block b covers elements b, b + gridDim.x*256, ... per thread.
template <int OP>
__global__ __launch_bounds__(256) void reduce1_kernel(int n, const double *__restrict__ x,
                                                      double *__restrict__ part)
{
    __shared__ double ws[4];
    double v = red_init<OP>();
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        v = red_combine<OP>(v, red_load<OP>(x[i]));
    for (int o = 32; o > 0; o >>= 1) v = red_combine<OP>(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.x] = red_combine<OP>(red_combine<OP>(red_combine<OP>(ws[0], ws[1]), ws[2]), ws[3]);
}
// Stage 2: one block folds the stage-1 results; out = alpha * result.
template <int OP>
__global__ __launch_bounds__(256) void reduce2_kernel(int nb, const double *__restrict__ part,
                                                      double alpha, double *out)
{
    __shared__ double ws[4];
    double v = red_init<OP>();
    for (int i = threadIdx.x; i < nb; i += 256) v = red_combine<OP>(v, part[i]);
    for (int o = 32; o > 0; o >>= 1) v = red_combine<OP>(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        *out = alpha * red_combine<OP>(red_combine<OP>(red_combine<OP>(ws[0], ws[1]), ws[2]), ws[3]);
}

// First index of max |x| (IDAMAX) with the signed value there.
__device__ __forceinline__ void amax_better(double &v, int &i, double v2, int i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}
__global__ __launch_bounds__(256) void iamax1_kernel(int n, const double *__restrict__ x,
                                                     double *__restrict__ pv, int *__restrict__ pi)
{
    __shared__ double wv[4];
    __shared__ int wi[4];
    double v = -1.0;
    int idx = INT_MAX;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        amax_better(v, idx, fabs(x[i]), i);
    for (int o = 32; o > 0; o >>= 1) amax_better(v, idx, __shfl_xor(v, o), __shfl_xor(idx, o));
    if ((threadIdx.x & 63) == 0) {
        wv[threadIdx.x >> 6] = v;
        wi[threadIdx.x >> 6] = idx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) amax_better(v, idx, wv[w], wi[w]);
        pv[blockIdx.x] = v;
        pi[blockIdx.x] = idx;
    }
}
// out_v = x[j] (signed), out_i = j (0-based).
__global__ __launch_bounds__(64) void iamax2_kernel(int nb, const double *__restrict__ pv,
                                                    const int *__restrict__ pi,
                                                    const double *__restrict__ x, double *out_v,
                                                    int *out_i)
{
    double v = -1.0;
    int idx = INT_MAX;
    for (int b = threadIdx.x; b < nb; b += 64) amax_better(v, idx, pv[b], pi[b]);
    for (int o = 32; o > 0; o >>= 1) amax_better(v, idx, __shfl_xor(v, o), __shfl_xor(idx, o));
    if (threadIdx.x == 0) {
        if (idx == INT_MAX) idx = 0;  // no comparable element (all NaN)
        *out_i = idx;
        *out_v = x[idx];
    }
}

// ---- estimator vectors (pdlacon.f) -----------------------------------------
__global__ void fill_kernel(int n, double *x, double v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void unit_kernel(int n, double *x, int j)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = (i == j) ? 1.0 : 0.0;
}
// x(k) = (-1)^(k+1) (1 + (k-1)/(n-1)), k = 1..n (pdlacon.f, label 120)
__global__ void altsgn_kernel(int n, double *x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = ((i + 1) % 2 == 0 ? -1.0 : 1.0) * (1.0 + (double)i / (double)(n - 1));
}
// compare: *changed = 1 if some NINT(SIGN(1, x)) differs from isgn; else
// x = SIGN(1, x), isgn = NINT(x).
__global__ void sign_kernel(int n, double *x, int *isgn, int compare, int *changed)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double s = copysign(1.0, x[i]);
    if (compare) {
        if ((int)s != isgn[i]) atomicOr(changed, 1);
    } else {
        x[i] = s;
        isgn[i] = (int)s;
    }
}
__global__ void copy_kernel(int n, const double *__restrict__ x, double *__restrict__ y)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[i];
}
__global__ void mul_kernel(int n, const double *__restrict__ w, double *__restrict__ x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = w[i] * x[i];
}
__global__ void axpy1_kernel(int n, const double *__restrict__ r, double *__restrict__ x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] = x[i] + r[i];
}

// ---- PDGERFS residual and |b| + |op(A)||x| ---------------------------------
// 'N': stage 1, block (x: 256 rows, y: 512-column slice) writes
// part[slice][0][i] = sum A(i, j) x(j), part[slice][1][i] = sum |A(i, j)| |x(j)|.
__global__ __launch_bounds__(256) void rfs_partial_n_kernel(int n, const double *__restrict__ A,
                                                            int lda, const double *__restrict__ x,
                                                            double *__restrict__ part)
{
    __shared__ double xs[RF_KC];
    const int c0 = blockIdx.y * RF_KC, kc = min(RF_KC, n - c0);
    for (int k = threadIdx.x; k < kc; k += 256) xs[k] = x[c0 + k];
    __syncthreads();
    const int i = blockIdx.x * RF_ROWS + threadIdx.x;
    if (i >= n) return;
    const double *a = A + i + (size_t)c0 * lda;
    double acc = 0.0, wacc = 0.0;
#pragma unroll 8
    for (int k = 0; k < kc; k++) {
        const double v = a[(size_t)k * lda], xv = xs[k];
        acc = fma(v, xv, acc);
        wacc = fma(fabs(v), fabs(xv), wacc);
    }
    double *P = part + (size_t)blockIdx.y * 2 * n;
    P[i] = acc;
    P[n + i] = wacc;
}

// 'T': stage 1, block (x: 64 columns, y: 1024-row slice); wave w takes
// columns 16w .. 16w+15, lanes read down each column.
__global__ __launch_bounds__(256) void rfs_partial_t_kernel(int n, const double *__restrict__ A,
                                                            int lda, const double *__restrict__ x,
                                                            double *__restrict__ part)
{
    __shared__ double xs[RF_TROWS];
    const int r0 = blockIdx.y * RF_TROWS, rs = min(RF_TROWS, n - r0);
    for (int k = threadIdx.x; k < rs; k += 256) xs[k] = x[r0 + k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double *P = part + (size_t)blockIdx.y * 2 * n;
    for (int cc = 0; cc < RF_TCOLS / 4; cc++) {
        const int col = blockIdx.x * RF_TCOLS + wave * (RF_TCOLS / 4) + cc;
        if (col >= n) break;
        const double *a = A + (size_t)col * lda + r0;
        double acc = 0.0, wacc = 0.0;
#pragma unroll 4
        for (int r = lane; r < rs; r += 64) {
            const double v = a[r], xv = xs[r];
            acc = fma(v, xv, acc);
            wacc = fma(fabs(v), fabs(xv), wacc);
        }
        acc = wave_sum(acc);
        wacc = wave_sum(wacc);
        if (lane == 0) {
            P[col] = acc;
            P[n + col] = wacc;
        }
    }
}

// Stage 2: r = b - sum, w = |b| + sum |.|, S = max_i of the componentwise
// ratio (pdgerfs.f:524-541) into *smax (bit-pattern max).
__global__ __launch_bounds__(256) void rfs_reduce_kernel(int n, int slices, const double *__restrict__ b,
                                                         const double *__restrict__ part,
                                                         double *__restrict__ r,
                                                         double *__restrict__ w, double safe1,
                                                         double safe2, unsigned long long *smax)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        double s1 = 0.0, s2 = 0.0;
        for (int c = 0; c < slices; c++) {
            s1 += part[(size_t)c * 2 * n + i];
            s2 += part[(size_t)c * 2 * n + n + i];
        }
        const double bi = b[i];
        const double ri = bi - s1, wi = fabs(bi) + s2;
        r[i] = ri;
        w[i] = wi;
        s = wi > safe2 ? fabs(ri) / wi : (fabs(ri) + safe1) / (wi + safe1);
    }
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(smax, s);
}

// w = |r| + nz eps w (+ safe1 where w <= safe2), pdgerfs.f:561-572.
// x[i] in [-1, 1): an integer hash of i (the power steps' fixed start).
__global__ void hash_vector_kernel(int n, double *x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t v = (uint32_t)i * 2654435761u + 0x9E3779B9u;
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    x[i] = (double)v / 2147483648.0 - 1.0;
}

// Stage 1 of x . y (fixed order, then reduce2_kernel<RED_SUM>).
__global__ __launch_bounds__(256) void dot1_kernel(int n, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   double *__restrict__ part)
{
    __shared__ double ws[4];
    double v = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) v = fma(x[i], y[i], v);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

__global__ void scal_kernel(int n, double alpha, double *x)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] *= alpha;
}

// ---- the selector's fused feature passes -----------------------------------
// Row pass (the block layout of rfs_partial_n_kernel: 256 rows x a 512-column
// slice per block): per row and slice the sum of |a| and the product A v
// (part[slice][0 / 1][row]), per block the largest |a| (bmax) and the sum of
// squares as a (scale, sum) pair of LAPACK DLASSQ's kind (bssq[2 block]).
__global__ __launch_bounds__(256) void feat_rows_kernel(int n, const double *__restrict__ A, int lda,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ part,
                                                        double *__restrict__ bmax,
                                                        double *__restrict__ bssq)
{
    __shared__ double xs[RF_KC];
    __shared__ double ws[3][4];
    const int c0 = blockIdx.y * RF_KC, kc = min(RF_KC, n - c0);
    for (int k = threadIdx.x; k < kc; k += 256) xs[k] = x[c0 + k];
    __syncthreads();
    const int i = blockIdx.x * RF_ROWS + threadIdx.x;
    double asum = 0.0, acc = 0.0, amax = 0.0, scale = 0.0, ssq = 1.0;
    if (i < n) {
        const double *a = A + i + (size_t)c0 * lda;
#pragma unroll 8
        for (int k = 0; k < kc; k++) {
            const double v = a[(size_t)k * lda], av = fabs(v);
            asum += av;
            acc = fma(v, xs[k], acc);
            amax = fmax(amax, av);
            if (av > 0.0) {
                if (scale < av) {
                    const double r = scale / av;
                    ssq = 1.0 + ssq * r * r;
                    scale = av;
                } else {
                    const double r = av / scale;
                    ssq = fma(r, r, ssq);
                }
            }
        }
        double *P = part + (size_t)blockIdx.y * 2 * n;
        P[i] = asum;
        P[n + i] = acc;
    }
    // block maximum and (scale, ssq) pair, folded in a fixed order
    for (int o = 32; o > 0; o >>= 1) {
        const double m2 = __shfl_xor(amax, o), s2 = __shfl_xor(scale, o), q2 = __shfl_xor(ssq, o);
        amax = fmax(amax, m2);
        const double sm = fmax(scale, s2);
        if (sm > 0.0) {
            const double r1 = scale / sm, r2 = s2 / sm;
            ssq = (scale > 0.0 ? ssq * r1 * r1 : 0.0) + (s2 > 0.0 ? q2 * r2 * r2 : 0.0);
        }
        scale = sm;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        ws[0][wave] = amax;
        ws[1][wave] = scale;
        ws[2][wave] = ssq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = ws[0][0], sc = ws[1][0], q = ws[2][0];
        for (int w = 1; w < 4; w++) {
            m = fmax(m, ws[0][w]);
            const double s2 = ws[1][w], sm = fmax(sc, s2);
            if (sm > 0.0) {
                const double r1 = sc / sm, r2 = s2 / sm;
                q = (sc > 0.0 ? q * r1 * r1 : 0.0) + (s2 > 0.0 ? ws[2][w] * r2 * r2 : 0.0);
            }
            sc = sm;
        }
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        bmax[b] = m;
        bssq[2 * b] = sc;
        bssq[2 * b + 1] = q;
    }
}

// Column pass (the layout of rfs_partial_t_kernel: 64 columns x a 1024-row
// slice per block, a wave reads down 16 columns): per column and slice the
// sum of |a| and A^T y (part[slice][0 / 1][col]).
__global__ __launch_bounds__(256) void feat_cols_kernel(int n, const double *__restrict__ A, int lda,
                                                        const double *__restrict__ y,
                                                        double *__restrict__ part)
{
    __shared__ double ys[RF_TROWS];
    const int r0 = blockIdx.y * RF_TROWS, rs = min(RF_TROWS, n - r0);
    for (int k = threadIdx.x; k < rs; k += 256) ys[k] = y[r0 + k];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double *P = part + (size_t)blockIdx.y * 2 * n;
    for (int cc = 0; cc < RF_TCOLS / 4; cc++) {
        const int col = blockIdx.x * RF_TCOLS + wave * (RF_TCOLS / 4) + cc;
        if (col >= n) break;
        const double *a = A + (size_t)col * lda + r0;
        double asum = 0.0, acc = 0.0;
#pragma unroll 4
        for (int r = lane; r < rs; r += 64) {
            const double v = a[r];
            asum += fabs(v);
            acc = fma(v, ys[r], acc);
        }
        asum = wave_sum(asum);
        acc = wave_sum(acc);
        if (lane == 0) {
            P[col] = asum;
            P[n + col] = acc;
        }
    }
}

// y[i] = sum over slices of part[slice][half][i] (fixed order).
__global__ __launch_bounds__(256) void slice_sum2_kernel(int n, int slices, int half,
                                                         const double *__restrict__ part,
                                                         double *__restrict__ y)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < slices; c++) s += part[(size_t)c * 2 * n + (size_t)half * n + i];
    y[i] = s;
}

// Fold of the per-block (scale, ssq) pairs: res = scale sqrt(ssq) (one block).
__global__ __launch_bounds__(64) void ssq_fold_kernel(int nb, const double *__restrict__ bssq,
                                                      double *res)
{
    if (threadIdx.x != 0) return;
    double sc = 0.0, q = 1.0;
    for (int b = 0; b < nb; b++) {
        const double s2 = bssq[2 * b], sm = fmax(sc, s2);
        if (sm > 0.0) {
            const double r1 = sc / sm, r2 = s2 / sm;
            q = (sc > 0.0 ? q * r1 * r1 : 0.0) + (s2 > 0.0 ? bssq[2 * b + 1] * r2 * r2 : 0.0);
        }
        sc = sm;
    }
    *res = sc * sqrt(q);
}

__global__ void ferr_weights_kernel(int n, const double *__restrict__ r, double *__restrict__ w,
                                    double nzeps, double safe1, double safe2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double wi = w[i];
    w[i] = wi > safe2 ? fabs(r[i]) + nzeps * wi : fabs(r[i]) + nzeps * wi + safe1;
}

// ---- selector features: sampled principal submatrices ---------------------
// Sample t (blockIdx.y) is A(idx_t, idx_t), k x k, copied twice (orig and
// work, ld k); its right-hand side is the alternating-sign vector of
// pdlacon.f's label 120.
__global__ void gather_sample_kernel(const double *__restrict__ A, int lda, int k,
                                     const int *__restrict__ idx, double *__restrict__ orig,
                                     double *__restrict__ work, double *__restrict__ rhs)
{
    const int t = blockIdx.y;
    const int *ix = idx + (size_t)t * k;
    const size_t kk = (size_t)k * k;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < kk; e += (size_t)gridDim.x * 256) {
        const int i = (int)(e % k), j = (int)(e / k);
        const double v = A[(size_t)ix[i] + (size_t)ix[j] * lda];
        orig[t * kk + e] = v;
        work[t * kk + e] = v;
    }
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < k; i += 256)
            rhs[(size_t)t * k + i] = ((i + 1) % 2 == 0 ? -1.0 : 1.0) *
                                     (1.0 + (k > 1 ? (double)i / (double)(k - 1) : 0.0));
}

// Per sample (one block each): max |A_s|, ||A_s||_1, max |L\U|, max and min
// |u_ii|, sum |x| (x = A_s^{-1} x_alt) into st[6 t ..].
__global__ __launch_bounds__(256) void sample_stats_kernel(int k, const double *__restrict__ orig,
                                                           const double *__restrict__ lu,
                                                           const double *__restrict__ x,
                                                           double *__restrict__ st)
{
    __shared__ double r[6][256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const size_t kk = (size_t)k * k;
    const double *a = orig + t * kk, *f = lu + t * kk, *xv = x + (size_t)t * k;
    double mA = 0.0, n1 = 0.0, mLU = 0.0, dmax = 0.0, dmin = INFINITY, sx = 0.0;
    for (size_t e = tid; e < kk; e += 256) {
        mA = fmax(mA, fabs(a[e]));
        mLU = fmax(mLU, fabs(f[e]));
    }
    for (int j = tid; j < k; j += 256) {
        double c = 0.0;
        for (int i = 0; i < k; i++) c += fabs(a[(size_t)i + (size_t)j * k]);
        n1 = fmax(n1, c);
        const double d = fabs(f[(size_t)j * (k + 1)]);
        dmax = fmax(dmax, d);
        dmin = fmin(dmin, d);
        sx += fabs(xv[j]);
    }
    r[0][tid] = mA;
    r[1][tid] = n1;
    r[2][tid] = mLU;
    r[3][tid] = dmax;
    r[4][tid] = dmin;
    r[5][tid] = sx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            r[0][tid] = fmax(r[0][tid], r[0][tid + w]);
            r[1][tid] = fmax(r[1][tid], r[1][tid + w]);
            r[2][tid] = fmax(r[2][tid], r[2][tid + w]);
            r[3][tid] = fmax(r[3][tid], r[3][tid + w]);
            r[4][tid] = fmin(r[4][tid], r[4][tid + w]);
            r[5][tid] += r[5][tid + w];
        }
        __syncthreads();
    }
    if (tid < 6) st[6 * t + tid] = r[tid][0];
}

// ---- PDGEEQU / PDLAQGE / PDGESVX ------------------------------------------
// r_i = max_j |a_ij| (bit-pattern max into rbits, zero on entry)
__global__ __launch_bounds__(256) void rowmax_kernel(int m, int n, const double *__restrict__ A,
                                                     int lda, unsigned long long *rbits)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double v = 0.0;
    for (int j = blockIdx.y; j < n; j += gridDim.y) v = fmax(v, fabs(A[i + (size_t)j * lda]));
    atomic_max_nonneg(rbits + i, v);
}
// c_j = max_i |a_ij| r_i, one wave per column
__global__ __launch_bounds__(256) void colmax_scaled_kernel(int m, int n, const double *__restrict__ A,
                                                            int lda, const double *__restrict__ r,
                                                            double *__restrict__ c)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= n) return;
    const double *a = A + (size_t)j * lda;
    double v = 0.0;
    for (int i = lane; i < m; i += 64) v = fmax(v, fabs(a[i]) * r[i]);
    v = wave_max(v);
    if (lane == 0) c[j] = v;
}
// v = 1 / min(max(v, smlnum), bignum)
__global__ void recip_clamp_kernel(int n, double *v, double smlnum, double bignum)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = 1.0 / fmin(fmax(v[i], smlnum), bignum);
}
// *idx = min { i : v[i] == 0 } (INT_MAX on entry)
__global__ void first_zero_kernel(int n, const double *v, int *idx)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n && v[i] == 0.0) atomicMin(idx, i);
}
// A = diag(r)^mr A diag(c)^mc, the product order of pdlaqge.f:226-258
__global__ void laqge_kernel(int m, int n, double *A, int lda, const double *__restrict__ r,
                             const double *__restrict__ c, int mode)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        double *a = A + i + (size_t)j * lda;
        if (mode == 1) *a = c[j] * *a;             // 'C'
        else if (mode == 2) *a = r[i] * *a;        // 'R'
        else *a = c[j] * r[i] * *a;                // 'B'
    }
}
// B(i, :) = v_i B(i, :)
__global__ void scale_rows_kernel(int n, int nrhs, const double *__restrict__ v, double *B, int ldb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < nrhs; j++) B[i + (size_t)j * ldb] = v[i] * B[i + (size_t)j * ldb];
}

unsigned blocks(long long n) { return (unsigned)std::max<long long>(1, (n + 255) / 256); }

}  // namespace

// ============================================================================
// Device workspace of the routines in this file
// ============================================================================
struct AuxWork {
    Handle &h;
    int n;
    double *x, *v, *r, *w, *part;
    int *isgn;
    double *red_v;           // [RED_BLOCKS] stage-1 values
    int *red_i;              // [RED_BLOCKS] stage-1 indices
    double *res;             // [8] scalar results
    int *ires;               // [8]
    unsigned long long *ures;  // [8]

    AuxWork(Handle &hh, int nn, size_t part_doubles) : h(hh), n(nn)
    {
        const size_t vn = ((size_t)std::max(n, 1) + 63) / 64 * 64;
        const size_t bytes = (4 * vn + part_doubles + RED_BLOCKS + 8) * 8 + (vn + RED_BLOCKS + 8) * 4 +
                             8 * 8 + 1024;
        char *p = static_cast<char *>(h.buffer(BUF_AUX, bytes));
        double *d = reinterpret_cast<double *>(p);
        x = d;
        v = d + vn;
        r = d + 2 * vn;
        w = d + 3 * vn;
        part = d + 4 * vn;
        red_v = part + part_doubles;
        res = red_v + RED_BLOCKS;
        ures = reinterpret_cast<unsigned long long *>(res + 8);
        isgn = reinterpret_cast<int *>(ures + 8);
        red_i = isgn + vn;
        ires = red_i + RED_BLOCKS;
    }
    hipStream_t st() const { return h.stream; }
    void fetch(void *dst, const void *src, size_t bytes)
    {
        SLA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st()));
        SLA_HIP(hipStreamSynchronize(st()));
    }
    int red_blocks(int len) const { return std::min(RED_BLOCKS, std::max(1, ceil_div(len, 1024))); }
    // alpha * sum |y| (PDASUM) into res[slot]
    void asum(const double *y, int len, int slot, double alpha = 1.0)
    {
        const int nb = red_blocks(len);
        hipLaunchKernelGGL(reduce1_kernel<RED_SUM_ABS>, dim3(nb), dim3(256), 0, st(), len, y, red_v);
        hipLaunchKernelGGL(reduce2_kernel<RED_SUM_ABS>, dim3(1), dim3(256), 0, st(), nb, red_v, alpha,
                           res + slot);
        SLA_LAUNCH_CHECK();
    }
    void maxabs(const double *y, int len, int slot)
    {
        const int nb = red_blocks(len);
        hipLaunchKernelGGL(reduce1_kernel<RED_MAX_ABS>, dim3(nb), dim3(256), 0, st(), len, y, red_v);
        hipLaunchKernelGGL(reduce2_kernel<RED_MAX_ABS>, dim3(1), dim3(256), 0, st(), nb, red_v, 1.0,
                           res + slot);
        SLA_LAUNCH_CHECK();
    }
    void minabs(const double *y, int len, int slot)
    {
        const int nb = red_blocks(len);
        hipLaunchKernelGGL(reduce1_kernel<RED_MIN_ABS>, dim3(nb), dim3(256), 0, st(), len, y, red_v);
        hipLaunchKernelGGL(reduce2_kernel<RED_MIN_ABS>, dim3(1), dim3(256), 0, st(), nb, red_v, 1.0,
                           res + slot);
        SLA_LAUNCH_CHECK();
    }
    // IDAMAX: res[slot] = y[j] (signed), ires[slot] = j (0-based)
    void iamax(const double *y, int len, int slot)
    {
        const int nb = red_blocks(len);
        hipLaunchKernelGGL(iamax1_kernel, dim3(nb), dim3(256), 0, st(), len, y, red_v, red_i);
        hipLaunchKernelGGL(iamax2_kernel, dim3(1), dim3(64), 0, st(), nb, red_v, red_i, y, res + slot,
                           ires + slot);
        SLA_LAUNCH_CHECK();
    }
};

// ============================================================================
// PDLANGE
// ============================================================================
double dlange_device(Handle &h, char norm, int m, int n, const double *dA, int lda)
{
    if (std::min(m, n) <= 0) return 0.0;  // pdlange.f:197-199
    const char c = (char)std::toupper((unsigned char)norm);
    if (c == 'M' || c == 'F' || c == 'E') {
        const size_t gx = ceil_div(m, 256);
        const size_t gy = std::min<size_t>((size_t)n, std::max<size_t>(1, 4096 / gx));
        AuxWork a(h, 1, gx * gy);
        fill_u32(reinterpret_cast<unsigned *>(a.ures), 2, 0u, h.stream);
        hipLaunchKernelGGL(lange_max_kernel, dim3(gx, gy), dim3(256), 0, h.stream, m, n, dA, lda,
                           a.ures);
        SLA_LAUNCH_CHECK();
        if (c == 'M') {
            unsigned long long bits = 0;
            a.fetch(&bits, a.ures, 8);
            double v;
            std::memcpy(&v, &bits, 8);
            return v;
        }
        hipLaunchKernelGGL(lange_ssq_kernel, dim3(gx, gy), dim3(256), 0, h.stream, m, n, dA, lda,
                           a.ures, a.part);
        SLA_LAUNCH_CHECK();
        hipLaunchKernelGGL(reduce2_kernel<RED_SUM>, dim3(1), dim3(256), 0, h.stream, (int)(gx * gy),
                           a.part, 1.0, a.res);
        SLA_LAUNCH_CHECK();
        double ssq = 0.0;
        unsigned long long bits = 0;
        a.fetch(&ssq, a.res, 8);
        a.fetch(&bits, a.ures, 8);
        double scale;
        std::memcpy(&scale, &bits, 8);
        return scale * std::sqrt(ssq);
    }
    if (c == 'O' || c == '1') {
        AuxWork a(h, n, 0);
        hipLaunchKernelGGL(lange_colsum_kernel, dim3(ceil_div(n, 64)), dim3(64), 0, h.stream, m, n, dA,
                           lda, a.x);
        SLA_LAUNCH_CHECK();
        a.maxabs(a.x, n, 0);
        double v = 0.0;
        a.fetch(&v, a.res, 8);
        return v;
    }
    if (c == 'I') {
        AuxWork a(h, m, 0);
        hipLaunchKernelGGL(lange_rowsum_kernel, dim3(ceil_div(m, 64)), dim3(64), 0, h.stream, m, n, dA,
                           lda, a.x);
        SLA_LAUNCH_CHECK();
        a.maxabs(a.x, m, 0);
        double v = 0.0;
        a.fetch(&v, a.res, 8);
        return v;
    }
    throw std::invalid_argument(std::string("dlange: unknown norm '") + norm + "'");
}

// ============================================================================
// PDLACON: estimate ||B||_1 for the operator B given by apply(1) (x <- B x)
// and apply(2) (x <- B^T x) acting on a.x.
//
// pdlacon.f:188-189 set EST = 0 and ESTWORK(1) = EST on every entry of the
// reverse-communication routine, so no estimate survives a return to the
// caller: label 70 compares with ESTOLD = 0 and label 140 compares TEMP
// with 0.  The value the reference returns is therefore always label 140's
// TEMP = 2 ||B x_alt||_1 / (3n) for the alternating-sign vector x_alt
// (or 0 when TEMP is not positive); the steps before it only decide how many
// products are formed.  reference_result computes exactly that value with
// one product.  With lapack_steps the loop of LAPACK's DLACON runs instead
// (jump labels 20 .. 140, ITMAX = 5, the estimate kept across steps), which
// is never smaller and usually much closer to ||B||_1.
// ============================================================================
namespace {

template <typename Apply>
double lacon(AuxWork &a, Apply apply, bool lapack_steps)
{
    const int n = a.n;
    const hipStream_t s = a.st();
    const unsigned g = blocks(n);
    constexpr int ITMAX = 5;
    if (n == 1) {  // label 20, N = 1: EST = |B (1/n)|
        hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, a.x, 1.0);
        apply(1);
        double x0 = 0.0;
        a.fetch(&x0, a.x, 8);
        return std::fabs(x0);
    }
    if (!lapack_steps) {  // reference_result: label 120 then label 140
        hipLaunchKernelGGL(altsgn_kernel, dim3(g), dim3(256), 0, s, n, a.x);
        apply(1);
        a.asum(a.x, n, 2);
        double temp = 0.0;
        a.fetch(&temp, a.res + 2, 8);
        temp = 2.0 * (temp / (double)(3 * n));
        return temp > 0.0 ? temp : 0.0;
    }
    hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, n, a.x, 1.0 / (double)n);
    apply(1);
    double est = 0.0;
    a.asum(a.x, n, 0);
    a.fetch(&est, a.res, 8);
    hipLaunchKernelGGL(sign_kernel, dim3(g), dim3(256), 0, s, n, a.x, a.isgn, 0, a.ires + 4);
    apply(2);
    // label 40
    a.iamax(a.x, n, 1);
    int j = 0;
    a.fetch(&j, a.ires + 1, 4);
    int iter = 2;
    for (;;) {
        // label 50: x = e_j
        hipLaunchKernelGGL(unit_kernel, dim3(g), dim3(256), 0, s, n, a.x, j);
        apply(1);
        // label 70
        hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, n, a.x, a.v);
        const double estold = est;
        a.asum(a.v, n, 0);
        fill_int(a.ires + 4, 1, 0, s);
        hipLaunchKernelGGL(sign_kernel, dim3(g), dim3(256), 0, s, n, a.x, a.isgn, 1, a.ires + 4);
        SLA_LAUNCH_CHECK();
        double r0 = 0.0;
        int changed = 0;
        a.fetch(&r0, a.res, 8);
        a.fetch(&changed, a.ires + 4, 4);
        est = r0;
        if (changed == 0 || est <= estold) break;  // to label 120
        hipLaunchKernelGGL(sign_kernel, dim3(g), dim3(256), 0, s, n, a.x, a.isgn, 0, a.ires + 4);
        apply(2);
        // label 110
        const int jlast = j;
        a.iamax(a.x, n, 1);
        double xj[2];
        a.fetch(&j, a.ires + 1, 4);
        a.fetch(&xj[0], a.res + 1, 8);                // XMAX (signed)
        a.fetch(&xj[1], a.x + jlast, 8);              // X(JLAST) (signed, as PDELGET)
        if (xj[1] != std::fabs(xj[0]) && iter < ITMAX) {
            iter++;
            continue;  // label 50
        }
        break;
    }
    // label 120: alternating-sign vector
    hipLaunchKernelGGL(altsgn_kernel, dim3(g), dim3(256), 0, s, n, a.x);
    apply(1);
    // label 140
    a.asum(a.x, n, 2);
    double temp = 0.0;
    a.fetch(&temp, a.res + 2, 8);
    temp = 2.0 * (temp / (double)(3 * n));
    if (temp > est) {
        hipLaunchKernelGGL(copy_kernel, dim3(g), dim3(256), 0, s, n, a.x, a.v);
        est = temp;
    }
    SLA_LAUNCH_CHECK();
    return est;
}

}  // namespace

// ============================================================================
// PDGECON
// ============================================================================
double dgecon_device(Handle &h, char norm, int n, const double *dLU, int lda, double anorm)
{
    const char c = (char)std::toupper((unsigned char)norm);
    const bool onenrm = (c == '1' || c == 'O');
    if (!onenrm && c != 'I') throw std::invalid_argument("dgecon: NORM must be '1', 'O' or 'I'");
    if (anorm < 0.0) throw std::invalid_argument("dgecon: ANORM < 0");
    // quick returns, pdgecon.f:287-297
    if (n == 0) return 1.0;
    if (anorm == 0.0) return 0.0;
    if (n == 1) return 1.0;
    AuxWork a(h, n, 0);
    const int kase1 = onenrm ? 1 : 2;
    fill_int(h.abort_flag, 1, 0, h.stream);
    fill_int(h.bad_pivot, 1, 0, h.stream);  // no permutation here: a stale flag must not fail
    TrsvWork tw = h.trsv_work(n);
    auto apply = [&](int kase) {
        if (kase == kase1) {  // x <- U^{-1} L^{-1} x
            trsv_persistent<double, double>(TrsvOp::L, n, 1, dLU, lda, a.x, n, tw, h.stream);
            trsv_persistent<double, double>(TrsvOp::U, n, 1, dLU, lda, a.x, n, tw, h.stream);
        } else {              // x <- L^{-T} U^{-T} x
            trsv_persistent<double, double>(TrsvOp::UT, n, 1, dLU, lda, a.x, n, tw, h.stream);
            trsv_persistent<double, double>(TrsvOp::LT, n, 1, dLU, lda, a.x, n, tw, h.stream);
        }
    };
    double ainvnm = 0.0;
    try {
        ainvnm = lacon(a, apply, h.lacon_lapack != 0);
    } catch (...) {
        h.trsv_epoch = tw.epoch;
        throw;
    }
    h.trsv_epoch = tw.epoch;
    check_solve_flags(h, "dgecon");
    // pdgecon.f:395-396 (SCALE = SL * SU = 1: PDLATRS does not scale)
    return ainvnm != 0.0 ? (1.0 / ainvnm) / anorm : 0.0;
}

// ============================================================================
// PDGERFS
// ============================================================================
void dgerfs_device(Handle &h, bool trans, int n, int nrhs, const double *dA, int lda,
                   const double *dAF, int ldaf, const int *dipiv, const double *dB, int ldb,
                   double *dX, int ldx, double *ferr, double *berr)
{
    if (nrhs <= 0) return;
    if (n <= 1) {  // pdgerfs.f:440-446
        for (int k = 0; k < nrhs; k++) ferr[k] = berr[k] = 0.0;
        return;
    }
    constexpr int ITMAX = 5;
    const double eps = EPS64, safmin = DBL_MIN;
    const int nz = n + 1;
    const double safe1 = nz * safmin, safe2 = safe1 / eps;
    const int slices = trans ? ceil_div(n, RF_TROWS) : ceil_div(n, RF_KC);
    AuxWork a(h, n, (size_t)slices * 2 * n);
    fill_int(h.bad_pivot, 1, 0, h.stream);
    fill_int(h.abort_flag, 1, 0, h.stream);
    perm_from_ipiv(h, n, dipiv);
    const hipStream_t s = h.stream;
    const unsigned g = blocks(n);

    // r = b - op(A) x, w = |b| + |op(A)||x|, S into res[3] / ures[0]
    auto residual = [&](const double *b, const double *x) -> double {
        fill_u32(reinterpret_cast<unsigned *>(a.ures), 2, 0u, s);
        if (!trans)
            hipLaunchKernelGGL(rfs_partial_n_kernel, dim3(ceil_div(n, RF_ROWS), slices), dim3(256), 0, s,
                               n, dA, lda, x, a.part);
        else
            hipLaunchKernelGGL(rfs_partial_t_kernel, dim3(ceil_div(n, RF_TCOLS), slices), dim3(256), 0,
                               s, n, dA, lda, x, a.part);
        hipLaunchKernelGGL(rfs_reduce_kernel, dim3(g), dim3(256), 0, s, n, slices, b, a.part, a.r, a.w,
                           safe1, safe2, a.ures);
        SLA_LAUNCH_CHECK();
        unsigned long long bits = 0;
        a.fetch(&bits, a.ures, 8);
        double sv;
        std::memcpy(&sv, &bits, 8);
        return sv;
    };

    for (int k = 0; k < nrhs; k++) {
        const double *b = dB + (size_t)k * ldb;
        double *x = dX + (size_t)k * ldx;
        int count = 1;
        double lstres = 3.0;
        double sv;
        for (;;) {  // label 20
            sv = residual(b, x);
            berr[k] = sv;
            if (sv > eps && 2.0 * sv <= lstres && count <= ITMAX) {
                getrs_perm_device(h, trans, n, 1, dAF, ldaf, a.r, n);
                hipLaunchKernelGGL(axpy1_kernel, dim3(g), dim3(256), 0, s, n, a.r, x);
                SLA_LAUNCH_CHECK();
                lstres = sv;
                count++;
                continue;
            }
            break;
        }
        // Forward error bound: ||op(A)^{-1} diag(w)||_inf by the estimator.
        hipLaunchKernelGGL(ferr_weights_kernel, dim3(g), dim3(256), 0, s, n, a.r, a.w,
                           (double)nz * eps, safe1, safe2);
        SLA_LAUNCH_CHECK();
        auto apply = [&](int kase) {
            if (kase == 1) {  // x <- diag(w) op(A)^{-T} x
                getrs_perm_device(h, !trans, n, 1, dAF, ldaf, a.x, n);
                hipLaunchKernelGGL(mul_kernel, dim3(g), dim3(256), 0, s, n, a.w, a.x);
            } else {          // x <- op(A)^{-1} diag(w) x
                hipLaunchKernelGGL(mul_kernel, dim3(g), dim3(256), 0, s, n, a.w, a.x);
                getrs_perm_device(h, trans, n, 1, dAF, ldaf, a.x, n);
            }
            SLA_LAUNCH_CHECK();
        };
        const double est = lacon(a, apply, h.lacon_lapack != 0);
        a.maxabs(x, n, 5);
        double xmax = 0.0;
        a.fetch(&xmax, a.res + 5, 8);
        if (xmax != 0.0) ferr[k] = est / xmax;  // pdgerfs.f:655-664
    }
    check_solve_flags(h, "dgerfs");
}

// ============================================================================
// PDGEEQU
// ============================================================================
int dgeequ_device(Handle &h, int m, int n, const double *dA, int lda, double *dR, double *dC,
                  double *rowcnd, double *colcnd, double *amax)
{
    const double smlnum = DBL_MIN, bignum = 1.0 / smlnum;
    if (m == 0 || n == 0) {  // pdgeequ.f quick return
        *rowcnd = 1.0;
        *colcnd = 1.0;
        *amax = 0.0;
        return 0;
    }
    const hipStream_t s = h.stream;
    AuxWork a(h, std::max(m, n), 0);
    fill_u32(reinterpret_cast<unsigned *>(dR), 2 * m, 0u, s);
    const unsigned gx = blocks(m), gy = (unsigned)std::min<long long>(n, std::max(1u, 4096u / gx));
    hipLaunchKernelGGL(rowmax_kernel, dim3(gx, gy), dim3(256), 0, s, m, n, dA, lda,
                       reinterpret_cast<unsigned long long *>(dR));
    SLA_LAUNCH_CHECK();
    a.minabs(dR, m, 0);
    a.maxabs(dR, m, 1);
    double mm[2];
    a.fetch(mm, a.res, sizeof mm);
    *amax = mm[1];
    if (mm[0] == 0.0) {
        fill_int(a.ires + 6, 1, INT_MAX, s);
        hipLaunchKernelGGL(first_zero_kernel, dim3(blocks(m)), dim3(256), 0, s, m, dR, a.ires + 6);
        int idx = 0;
        a.fetch(&idx, a.ires + 6, 4);
        return idx + 1;
    }
    hipLaunchKernelGGL(recip_clamp_kernel, dim3(blocks(m)), dim3(256), 0, s, m, dR, smlnum, bignum);
    *rowcnd = std::max(mm[0], smlnum) / std::min(mm[1], bignum);
    hipLaunchKernelGGL(colmax_scaled_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, m, n, dA, lda, dR,
                       dC);
    SLA_LAUNCH_CHECK();
    a.minabs(dC, n, 0);
    a.maxabs(dC, n, 1);
    a.fetch(mm, a.res, sizeof mm);
    if (mm[0] == 0.0) {
        fill_int(a.ires + 6, 1, INT_MAX, s);
        hipLaunchKernelGGL(first_zero_kernel, dim3(blocks(n)), dim3(256), 0, s, n, dC, a.ires + 6);
        int idx = 0;
        a.fetch(&idx, a.ires + 6, 4);
        return m + idx + 1;
    }
    hipLaunchKernelGGL(recip_clamp_kernel, dim3(blocks(n)), dim3(256), 0, s, n, dC, smlnum, bignum);
    SLA_LAUNCH_CHECK();
    *colcnd = std::max(mm[0], smlnum) / std::min(mm[1], bignum);
    SLA_HIP(hipStreamSynchronize(s));
    return 0;
}

// ============================================================================
// PDLAQGE (THRESH = 0.1; SMALL = SAFMIN / PRECISION, PRECISION = 2^-52)
// ============================================================================
char dlaqge_device(Handle &h, int m, int n, double *dA, int lda, const double *dR, const double *dC,
                   double rowcnd, double colcnd, double amax)
{
    const double thresh = 0.1, small = DBL_MIN / std::ldexp(1.0, -52), large = 1.0 / small;
    if (m <= 0 || n <= 0) return 'N';
    int mode;
    char equed;
    if (rowcnd >= thresh && amax >= small && amax <= large) {
        if (colcnd >= thresh) return 'N';
        mode = 1;
        equed = 'C';
    } else if (colcnd >= thresh) {
        mode = 2;
        equed = 'R';
    } else {
        mode = 3;
        equed = 'B';
    }
    const unsigned gx = blocks(m), gy = (unsigned)std::min<long long>(n, std::max(1u, 8192u / gx));
    hipLaunchKernelGGL(laqge_kernel, dim3(gx, gy), dim3(256), 0, h.stream, m, n, dA, lda, dR, dC, mode);
    SLA_LAUNCH_CHECK();
    SLA_HIP(hipStreamSynchronize(h.stream));
    return equed;
}

// ============================================================================
// PDGESVX (SRC/pdgesvx.f:624-737)
// ============================================================================
int dgesvx_device(Handle &h, char fact, bool trans, int n, int nrhs, double *dA, int lda, double *dAF,
                  int ldaf, int *dipiv, char *equed, double *dR, double *dC, double *dB, int ldb,
                  double *dX, int ldx, double *rcond, double *ferr, double *berr)
{
    const hipStream_t s = h.stream;
    const double smlnum = DBL_MIN, bignum = 1.0 / smlnum;
    const char f = (char)std::toupper((unsigned char)fact);
    bool rowequ = false, colequ = false;
    double rowcnd = 1.0, colcnd = 1.0, amax = 0.0;
    if (f == 'F') {
        const char e = (char)std::toupper((unsigned char)*equed);
        rowequ = e == 'R' || e == 'B';
        colequ = e == 'C' || e == 'B';
        AuxWork a(h, std::max(n, 1), 0);
        if (rowequ && n > 0) {  // pdgesvx.f:548-565
            a.minabs(dR, n, 0);
            a.maxabs(dR, n, 1);
            double mm[2];
            a.fetch(mm, a.res, sizeof mm);
            rowcnd = std::max(mm[0], smlnum) / std::min(mm[1], bignum);
        }
        if (colequ && n > 0) {
            a.minabs(dC, n, 0);
            a.maxabs(dC, n, 1);
            double mm[2];
            a.fetch(mm, a.res, sizeof mm);
            colcnd = std::max(mm[0], smlnum) / std::min(mm[1], bignum);
        }
    } else {
        *equed = 'N';
    }
    if (f == 'E') {
        if (dgeequ_device(h, n, n, dA, lda, dR, dC, &rowcnd, &colcnd, &amax) == 0) {
            *equed = dlaqge_device(h, n, n, dA, lda, dR, dC, rowcnd, colcnd, amax);
            rowequ = *equed == 'R' || *equed == 'B';
            colequ = *equed == 'C' || *equed == 'B';
        }
    }
    // scale B (pdgesvx.f:672-697)
    if (!trans && rowequ)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks(n)), dim3(256), 0, s, n, nrhs, dR, dB, ldb);
    else if (trans && colequ)
        hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks(n)), dim3(256), 0, s, n, nrhs, dC, dB, ldb);
    SLA_LAUNCH_CHECK();
    if (f == 'N' || f == 'E') {
        SLA_HIP(hipMemcpy2DAsync(dAF, (size_t)ldaf * 8, dA, (size_t)lda * 8, (size_t)n * 8, n,
                                 hipMemcpyDeviceToDevice, s));
        const int info = dgetrf_device(h, n, n, dAF, ldaf, dipiv);
        if (info != 0) {
            if (info > 0) *rcond = 0.0;
            return info;
        }
    }
    const char norm = trans ? 'I' : '1';
    const double anorm = dlange_device(h, norm, n, n, dA, lda);
    *rcond = dgecon_device(h, norm, n, dAF, ldaf, anorm);
    if (*rcond < EPS64) return n + 1;  // pdgesvx.f:715-718 (INFO = IA + N)
    SLA_HIP(hipMemcpy2DAsync(dX, (size_t)ldx * 8, dB, (size_t)ldb * 8, (size_t)n * 8, nrhs,
                             hipMemcpyDeviceToDevice, s));
    dgetrs_device(h, n, nrhs, dAF, ldaf, dipiv, dX, ldx, trans);
    dgerfs_device(h, trans, n, nrhs, dA, lda, dAF, ldaf, dipiv, dB, ldb, dX, ldx, ferr, berr);
    // undo the scaling of X and FERR (pdgesvx.f:737-780)
    if (!trans) {
        if (colequ) {
            hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks(n)), dim3(256), 0, s, n, nrhs, dC, dX, ldx);
            for (int j = 0; j < nrhs; j++) ferr[j] = ferr[j] / colcnd;
        }
    } else if (rowequ) {
        hipLaunchKernelGGL(scale_rows_kernel, dim3(blocks(n)), dim3(256), 0, s, n, nrhs, dR, dX, ldx);
        for (int j = 0; j < nrhs; j++) ferr[j] = ferr[j] / rowcnd;
    }
    SLA_LAUNCH_CHECK();
    SLA_HIP(hipStreamSynchronize(s));
    return 0;
}

// ============================================================================
// Features of the learned precision selector (12 floats):
//   0 log10 n
//   1 log10 ||A||_1     2 log10 ||A||_inf     3 log10 max |a|
//   4 log10 (||A||_F / sqrt n)
//   5 log10 (max / min row sum of |A|)   6 log10 (max / min column sum)
//   7 mean over samples of log10 (max |L\U| / max |A_s|)   (sampled growth)
//   8 max over samples of log10 (max |u_ii| / min |u_ii|)
//   9 mean over samples of log10 of the PDGECON-style estimate
//     ||A_s||_1 * 2 ||A_s^{-1} x_alt||_1 / (3k) (pdlacon.f label 140)
//  10 max over samples of the same
//  11 log10 (n / k)
// Samples: four principal k x k submatrices (k = min(n, 128); one when n <=
// 128) on index sets drawn by a fixed LCG, factored and solved in one
// batched launch.
// ============================================================================
namespace {
double lg(double x) { return std::log10(std::min(std::max(x, 1e-300), 1e300)); }
}  // namespace

// The selector's global features in four passes over A: a row pass (row
// sums of |A|, max |a|, the sum of squares, y1 = A v), a column pass (column
// sums of |A|, z1 = A^T y1), then y2 = A z1 and z2 = A^T y2 with z1 scaled
// to max |z1| = 1; lambda = (z1 . z2) / (z1 . z1) is the Rayleigh estimate
// of ||A||_2^2 of two power steps on A^T A started at the hash vector v.
struct FusedNorms {
    double n1, ni, nm, nf, rmin, cmin, lam;
};

static FusedNorms fused_norms(Handle &h, int n, const double *dA, int lda)
{
    const int sn = ceil_div(n, RF_KC), stn = ceil_div(n, RF_TROWS), rb = ceil_div(n, RF_ROWS);
    const int nblk = rb * sn;
    const size_t pd = (size_t)std::max(sn, stn) * 2 * n;
    AuxWork a(h, n, pd + 3 * (size_t)nblk);
    double *bmax = a.part + pd, *bssq = bmax + nblk;
    hipStream_t s = h.stream;
    const dim3 gs(ceil_div(n, 256));
    auto sum2 = [&](int slices, int half, double *out) {
        hipLaunchKernelGGL(slice_sum2_kernel, gs, dim3(256), 0, s, n, slices, half, a.part, out);
    };
    auto dot = [&](const double *x, const double *y, int slot) {
        const int nb = a.red_blocks(n);
        hipLaunchKernelGGL(dot1_kernel, dim3(nb), dim3(256), 0, s, n, x, y, a.red_v);
        hipLaunchKernelGGL(reduce2_kernel<RED_SUM>, dim3(1), dim3(256), 0, s, nb, a.red_v, 1.0,
                           a.res + slot);
    };
    hipLaunchKernelGGL(hash_vector_kernel, gs, dim3(256), 0, s, n, a.x);
    // pass 1: rows
    hipLaunchKernelGGL(feat_rows_kernel, dim3(rb, sn), dim3(256), 0, s, n, dA, lda, a.x, a.part, bmax,
                       bssq);
    sum2(sn, 0, a.w);  // row sums of |A|
    sum2(sn, 1, a.v);  // y1
    a.maxabs(a.w, n, 0);
    a.minabs(a.w, n, 1);
    a.maxabs(bmax, nblk, 2);
    hipLaunchKernelGGL(ssq_fold_kernel, dim3(1), dim3(64), 0, s, nblk, bssq, a.res + 3);
    // pass 2: columns
    hipLaunchKernelGGL(feat_cols_kernel, dim3(ceil_div(n, RF_TCOLS), stn), dim3(256), 0, s, n, dA,
                       lda, a.v, a.part);
    sum2(stn, 0, a.w);  // column sums of |A|
    sum2(stn, 1, a.r);  // z1
    a.maxabs(a.w, n, 4);
    a.minabs(a.w, n, 5);
    a.maxabs(a.r, n, 6);
    SLA_LAUNCH_CHECK();
    double r[7];
    a.fetch(r, a.res, sizeof r);
    FusedNorms out{r[4], r[0], r[2], r[3], r[1], r[5], 0.0};
    const double zmax = r[6];
    if (!(zmax > 0.0) || !std::isfinite(zmax)) return out;
    // passes 3 and 4: one more power step from z1 / max |z1|
    hipLaunchKernelGGL(scal_kernel, gs, dim3(256), 0, s, n, 1.0 / zmax, a.r);
    hipLaunchKernelGGL(rfs_partial_n_kernel, dim3(rb, sn), dim3(256), 0, s, n, dA, lda, a.r, a.part);
    sum2(sn, 0, a.w);  // y2
    hipLaunchKernelGGL(rfs_partial_t_kernel, dim3(ceil_div(n, RF_TCOLS), stn), dim3(256), 0, s, n, dA,
                       lda, a.w, a.part);
    sum2(stn, 0, a.x);  // z2
    dot(a.r, a.x, 0);
    dot(a.r, a.r, 1);
    SLA_LAUNCH_CHECK();
    double d[2];
    a.fetch(d, a.res, sizeof d);
    out.lam = d[1] > 0.0 ? d[0] / d[1] : 0.0;
    return out;
}

void selector_features_device(Handle &h, int n, const double *dA, int lda, float *f)
{
    if (n < 1) throw std::invalid_argument("selector features: n >= 1");
    f[0] = (float)std::log10((double)n);
    const FusedNorms fn = fused_norms(h, n, dA, lda);
    const double nf = fn.nf;
    f[1] = (float)lg(fn.n1);
    f[2] = (float)lg(fn.ni);
    f[3] = (float)lg(fn.nm);
    f[4] = (float)lg(nf / std::sqrt((double)n));
    f[5] = (float)(lg(fn.ni) - lg(fn.rmin));
    f[6] = (float)(lg(fn.n1) - lg(fn.cmin));
    const int k = std::min(n, 128), S = n > 128 ? 4 : 1;
    const size_t kk = (size_t)k * k;
    const size_t bytes = 2 * S * kk * 8 + S * (size_t)k * 8 + S * (size_t)k * 4 + S * 4 +
                         S * (size_t)k * 4 + 3 * S * sizeof(void *) + S * 6 * 8 + 1024;
    char *buf = static_cast<char *>(h.buffer(BUF_SELECT, bytes));
    double *orig = reinterpret_cast<double *>(buf);
    double *work = orig + S * kk;
    double *rhs = work + S * kk;
    double *stats = rhs + (size_t)S * k;
    int *piv = reinterpret_cast<int *>(stats + 6 * S);
    int *info = piv + (size_t)S * k;
    int *idx = info + S;
    void **ptrs = reinterpret_cast<void **>((reinterpret_cast<uintptr_t>(idx + (size_t)S * k) + 15) & ~uintptr_t(15));
    // index sets: partial Fisher-Yates with a fixed LCG, sorted
    std::vector<int> hidx((size_t)S * k);
    {
        std::vector<int> perm(n);
        for (int t = 0; t < S; t++) {
            for (int i = 0; i < n; i++) perm[i] = i;
            uint64_t st = 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1) + (uint64_t)n;
            for (int i = 0; i < k; i++) {
                st = st * 6364136223846793005ull + 1442695040888963407ull;
                const int j = i + (int)((st >> 33) % (uint64_t)(n - i));
                std::swap(perm[i], perm[j]);
            }
            std::sort(perm.begin(), perm.begin() + k);
            std::copy(perm.begin(), perm.begin() + k, hidx.begin() + (size_t)t * k);
        }
    }
    std::vector<void *> hp(3 * (size_t)S);
    for (int t = 0; t < S; t++) {
        hp[t] = work + t * kk;
        hp[S + t] = piv + (size_t)t * k;
        hp[2 * S + t] = rhs + (size_t)t * k;
    }
    SLA_HIP(hipMemcpyAsync(idx, hidx.data(), sizeof(int) * hidx.size(), hipMemcpyHostToDevice, h.stream));
    SLA_HIP(hipMemcpyAsync(ptrs, hp.data(), sizeof(void *) * hp.size(), hipMemcpyHostToDevice, h.stream));
    hipLaunchKernelGGL(gather_sample_kernel, dim3(std::max(1, (int)std::min<size_t>(64, (kk + 255) / 256)), S),
                       dim3(256), 0, h.stream, dA, lda, k, idx, orig, work, rhs);
    SLA_LAUNCH_CHECK();
    std::vector<int> nv(S, k), rv(S, 1);
    const int save_graph = h.batched_graph;
    h.batched_graph = 0;
    try {
        gesv_batched_device<double>(h, S, nv.data(), rv.data(), reinterpret_cast<double *const *>(ptrs),
                                    nv.data(), reinterpret_cast<int *const *>(ptrs + S),
                                    reinterpret_cast<double *const *>(ptrs + 2 * S), nv.data(), info);
    } catch (...) {
        h.batched_graph = save_graph;
        throw;
    }
    h.batched_graph = save_graph;
    hipLaunchKernelGGL(sample_stats_kernel, dim3(S), dim3(256), 0, h.stream, k, orig, work, rhs, stats);
    SLA_LAUNCH_CHECK();
    std::vector<double> hs(6 * (size_t)S);
    std::vector<int> hinfo(S);
    SLA_HIP(hipMemcpyAsync(hs.data(), stats, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * S, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    double g_mean = 0.0, d_max = -1e300, c_mean = 0.0, c_max = -1e300;
    for (int t = 0; t < S; t++) {
        const double *q = &hs[6 * (size_t)t];
        const double growth = lg(q[2]) - lg(q[0]);
        const double dr = lg(q[3]) - lg(q[4]);
        // a sample with an exactly zero pivot has no finite estimate: 1e20
        const double c = hinfo[t] != 0 ? 20.0 : std::min(20.0, lg(q[1] * 2.0 * q[5] / (3.0 * k)));
        g_mean += growth / S;
        d_max = std::max(d_max, std::min(20.0, dr));
        c_mean += c / S;
        c_max = std::max(c_max, c);
    }
    f[7] = (float)g_mean;
    f[8] = (float)d_max;
    f[9] = (float)c_mean;
    f[10] = (float)c_max;
    // the stable-rank fraction ||A||_F^2 / (n ||A||_2^2): a global view of
    // how the singular values spread, which the sampled submatrices of a
    // large matrix cannot see
    const double lam = fn.lam;
    f[11] = (float)(lam > 0.0 ? std::min(20.0, std::max(-20.0, lg(nf * nf / ((double)n * lam)))) : -20.0);
}

}  // namespace sla
