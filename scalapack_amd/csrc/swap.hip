// swap.hip -- row interchanges of the LU path (PDLASWP, SRC/pdlaswp.f:163-172;
// PDLAPIV/PDLAPV2 for the right-hand sides, SRC/pdlapv2.f:208-242).
//
// The reference applies the pivots of a panel one PDSWAP at a time, each a
// dswap_ of a row strided by LLD (PBLAS/SRC/PTOOLS/PB_CpswapNN.c:267).  In
// column-major HBM that is one cache line per element per swap.  Here the
// interchanges of a pivot block are first reduced to their net effect -- a
// list of (dst, src) row pairs, at most two per pivot -- by one small kernel,
// and every column tile then applies the whole list in one pass: all sources
// of a column are read into registers before any destination is written, so
// the result is the same as the sequential swaps, bit for bit.
#include "sla_common.h"
#include "sla_kernels.h"

namespace sla {

namespace {

constexpr int PAIRS_MAX_PIV = 512;

// One workgroup.  Slots 0..n-1 stand for rows k1..k2-1; slot n+i stands for
// pivot row ipiv[k1+i] outside [k1, k2) at its first occurrence (factorization
// pivots are >= their row, a caller's IPIV may hold any row).  The
// interchanges are replayed on slot contents (row ids), then each slot s
// emits the pair (row_of(s), content(s)).  A pivot row outside [0, nrows)
// raises *bad and the block emits no pairs.
__global__ __launch_bounds__(512) void laswp_pairs_kernel(const int *__restrict__ ipiv, int k1,
                                                          int k2, int nrows,
                                                          int *__restrict__ pairs,
                                                          int *__restrict__ count, int *bad)
{
    __shared__ int tgt[PAIRS_MAX_PIV];
    __shared__ int first[PAIRS_MAX_PIV];
    __shared__ int content[2 * PAIRS_MAX_PIV];
    __shared__ int s_bad;
    const int n = k2 - k1, tid = threadIdx.x;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = ipiv[k1 + i] - 1;
        tgt[i] = t;
        if (t < 0 || t >= nrows) s_bad = 1;
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) {
            *count = 0;
            atomicExch(bad, 1);
        }
        return;
    }
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = tgt[i];
        int f = -1;
        if (t < k1 || t >= k2) {
            f = i;
            for (int i2 = 0; i2 < i; i2++)
                if (tgt[i2] == t) { f = i2; break; }
        }
        first[i] = f;
    }
    __syncthreads();
    for (int s = tid; s < 2 * n; s += blockDim.x) {
        if (s < n) content[s] = k1 + s;
        else content[s] = (first[s - n] == s - n) ? tgt[s - n] : -1;
    }
    __syncthreads();
    if (tid == 0) {
        for (int i = 0; i < n; i++) {
            const int t = tgt[i];
            const int s = (t >= k1 && t < k2) ? (t - k1) : (n + first[i]);
            if (s != i) {
                const int x = content[i];
                content[i] = content[s];
                content[s] = x;
            }
        }
    }
    __syncthreads();
    for (int s = tid; s < 2 * n; s += blockDim.x) {
        int row;
        if (s < n) row = k1 + s;
        else row = (first[s - n] == s - n) ? tgt[s - n] : -1;
        pairs[2 * s] = row;
        pairs[2 * s + 1] = (row >= 0) ? content[s] : -1;
    }
    if (tid == 0) *count = 2 * n;
}

constexpr int AP_THREADS = 256, AP_COLS = 4, AP_PPT = 4;  // <= 1024 pairs

template <typename T>
__global__ __launch_bounds__(AP_THREADS) void laswp_apply_kernel(T *__restrict__ A, int lda,
                                                                  int roff, int ncols,
                                                                  const int *__restrict__ pairs,
                                                                  const int *__restrict__ count,
                                                                  int inverse)
{
    __shared__ int sp[2 * AP_THREADS * AP_PPT];
    const int cnt = *count, tid = threadIdx.x;
    // The inverse permutation (interchanges replayed backward) moves each
    // pair the other way: (dst, src) -> (src, dst).
    for (int i = tid; i < cnt; i += AP_THREADS) {
        const int d = pairs[2 * i], s = pairs[2 * i + 1];
        sp[2 * i] = inverse ? s : d;
        sp[2 * i + 1] = inverse ? d : s;
    }
    __syncthreads();
    const int c0 = blockIdx.x * AP_COLS;
    T v[AP_COLS][AP_PPT];
#pragma unroll
    for (int cc = 0; cc < AP_COLS; cc++) {
        const int c = c0 + cc;
#pragma unroll
        for (int q = 0; q < AP_PPT; q++) {
            const int idx = tid + AP_THREADS * q;
            if (c < ncols && idx < cnt) {
                const int dst = sp[2 * idx], src = sp[2 * idx + 1];
                if (dst >= 0 && dst != src) v[cc][q] = A[(size_t)(src - roff) + (size_t)c * lda];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < AP_COLS; cc++) {
        const int c = c0 + cc;
#pragma unroll
        for (int q = 0; q < AP_PPT; q++) {
            const int idx = tid + AP_THREADS * q;
            if (c < ncols && idx < cnt) {
                const int dst = sp[2 * idx], src = sp[2 * idx + 1];
                if (dst >= 0 && dst != src) A[(size_t)(dst - roff) + (size_t)c * lda] = v[cc][q];
            }
        }
    }
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_u32_kernel(unsigned *p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void iota_int_kernel(int *p, int n, int base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = base + i;
}

template <typename T>
__global__ void gather_rows_kernel(const T *__restrict__ src, int lds, T *__restrict__ dst,
                                   int ldd, const int *__restrict__ perm, int n, int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    for (int j = 0; j < nrhs; j++) dst[i + (size_t)j * ldd] = src[p + (size_t)j * lds];
}

template <typename T>
__global__ void scatter_rows_kernel(const T *__restrict__ src, int lds, T *__restrict__ dst,
                                    int ldd, const int *__restrict__ perm, int n, int nrhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    for (int j = 0; j < nrhs; j++) dst[p + (size_t)j * ldd] = src[i + (size_t)j * lds];
}

}  // namespace

void laswp_pairs(const int *ipiv, int k1, int k2, int nrows, int *pairs, int *count, int *bad,
                 hipStream_t s)
{
    if (k2 - k1 > PAIRS_MAX_PIV || k2 <= k1)
        throw std::logic_error("laswp_pairs: pivot block must hold 1..512 pivots");
    hipLaunchKernelGGL(laswp_pairs_kernel, dim3(1), dim3(512), 0, s, ipiv, k1, k2, nrows, pairs,
                       count, bad);
    SLA_LAUNCH_CHECK();
}

template <typename T>
void laswp_apply(T *A, int lda, int roff, int ncols, const int *pairs, const int *count,
                 int max_pairs, hipStream_t s, bool inverse)
{
    if (ncols <= 0) return;
    if (max_pairs > AP_THREADS * AP_PPT) throw std::logic_error("laswp_apply: too many pairs");
    hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3(ceil_div(ncols, AP_COLS)), dim3(AP_THREADS), 0,
                       s, A, lda, roff, ncols, pairs, count, inverse ? 1 : 0);
    SLA_LAUNCH_CHECK();
}

template void laswp_apply<double>(double *, int, int, int, const int *, const int *, int, hipStream_t, bool);
template void laswp_apply<float>(float *, int, int, int, const int *, const int *, int, hipStream_t, bool);
template void laswp_apply<int>(int *, int, int, int, const int *, const int *, int, hipStream_t, bool);

// Column interchanges (PDLASWP with ROWCOL = 'C', pdlaswp.f:186-203): one
// workgroup replays them in order, each one a contiguous column exchange.
__global__ __launch_bounds__(256) void laswp_cols_kernel(double *A, int lda, int nrows, int k1,
                                                         int k2, const int *ipiv, int forward)
{
    const int cnt = k2 - k1;
    for (int t = 0; t < cnt; t++) {
        const int i = forward ? k1 + t : k2 - 1 - t;
        const int p = ipiv[i] - 1;
        if (p != i) {
            double *ci = A + (size_t)i * lda, *cp = A + (size_t)p * lda;
            for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
                const double x = ci[r];
                ci[r] = cp[r];
                cp[r] = x;
            }
        }
        __syncthreads();
    }
}

void laswp_cols(double *A, int lda, int nrows, int k1, int k2, const int *ipiv, bool forward,
                hipStream_t s)
{
    if (k2 <= k1 || nrows <= 0) return;
    hipLaunchKernelGGL(laswp_cols_kernel, dim3(1), dim3(256), 0, s, A, lda, nrows, k1, k2, ipiv,
                       forward ? 1 : 0);
    SLA_LAUNCH_CHECK();
}

void fill_int(int *p, int n, int v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    SLA_LAUNCH_CHECK();
}
void fill_u32(unsigned *p, int n, unsigned v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    SLA_LAUNCH_CHECK();
}
void iota_int(int *p, int n, int base, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(iota_int_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, base);
    SLA_LAUNCH_CHECK();
}

template <typename T>
void gather_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n, int nrhs,
                 hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return;
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, src, lds,
                       dst, ldd, perm, n, nrhs);
    SLA_LAUNCH_CHECK();
}
template void gather_rows<double>(const double *, int, double *, int, const int *, int, int, hipStream_t);
template void gather_rows<float>(const float *, int, float *, int, const int *, int, int, hipStream_t);

template <typename T>
void scatter_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n, int nrhs,
                  hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return;
    hipLaunchKernelGGL(scatter_rows_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, src, lds,
                       dst, ldd, perm, n, nrhs);
    SLA_LAUNCH_CHECK();
}
template void scatter_rows<double>(const double *, int, double *, int, const int *, int, int, hipStream_t);

}  // namespace sla
