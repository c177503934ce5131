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

#include <climits>

namespace sla {

namespace {

constexpr int PAIRS_MAX_PIV = 512;
constexpr int PAIRS_HASH_BITS = 10, PAIRS_HASH = 1 << PAIRS_HASH_BITS;

// One workgroup.  Slots 0..n-1 stand for rows k1..k2-1; slot n+i stands for
// pivot row ipiv[k1+i] outside [k1, k2) at its first occurrence.  Step i
// exchanges the contents (row ids) of slot i and slot s_i; each slot x then
// emits the pair (row_of(x), content(x)).  A pivot row outside [0, nrows)
// raises *bad and the block emits no pairs.
//
// Factorization pivots are >= their row, so s_i >= i and slot i is final
// after step i.  Then the replay needs no sequential loop: with val(i) the
// content of slot i just before step i,
//   val(i)   = val(pred(i)), pred(i) = last step before i whose s is i,
//              or orig(i) when there is none (pointer jumping, log2 n rounds);
//   final(i) = val(i) if s_i = i, else val(l) for the last step l < i with
//              s_l = s_i, else orig(s_i);
//   final(x) for an outside slot = val(last step with s = x), else orig(x).
// A caller's IPIV may point backwards (any row): that block is replayed in
// order by one thread.
__global__ __launch_bounds__(512) void laswp_pairs_kernel(const int *__restrict__ ipiv, int k1,
                                                          int k2, int nrows,
                                                          int *__restrict__ pairs,
                                                          int *__restrict__ count, int *bad)
{
    __shared__ int tgt[PAIRS_MAX_PIV];
    __shared__ int first[PAIRS_MAX_PIV];
    __shared__ int slot[PAIRS_MAX_PIV];
    __shared__ int prev[PAIRS_MAX_PIV];
    __shared__ int jump[PAIRS_MAX_PIV];
    __shared__ int val[PAIRS_MAX_PIV];
    __shared__ int orig[2 * PAIRS_MAX_PIV];
    __shared__ int last[2 * PAIRS_MAX_PIV];
    __shared__ int cnt[2 * PAIRS_MAX_PIV];
    __shared__ int hkey[PAIRS_HASH], hval[PAIRS_HASH];
    __shared__ int s_bad, s_back;
    const int n = k2 - k1, tid = threadIdx.x;
    if (tid == 0) {
        s_bad = 0;
        s_back = 0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = ipiv[k1 + i] - 1;
        tgt[i] = t;
        if (t < 0 || t >= nrows) s_bad = 1;
    }
    for (int x = tid; x < PAIRS_HASH; x += blockDim.x) {
        hkey[x] = -1;
        hval[x] = INT_MAX;
    }
    __syncthreads();
    if (s_bad) {
        if (tid == 0) {
            *count = 0;
            atomicExch(bad, 1);
        }
        return;
    }
    // first occurrence of each outside row: open-addressing table (<= 512
    // keys in PAIRS_HASH entries), the lowest step per key by atomicMin
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = tgt[i];
        if (t < k1 || t >= k2) {
            int hx = (int)(((unsigned)t * 2654435761u) >> (32 - PAIRS_HASH_BITS));
            for (;;) {
                const int old = atomicCAS(&hkey[hx], -1, t);
                if (old == -1 || old == t) break;
                hx = (hx + 1) & (PAIRS_HASH - 1);
            }
            atomicMin(&hval[hx], i);
            first[i] = hx;  // the entry, resolved below
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = tgt[i];
        first[i] = (t < k1 || t >= k2) ? hval[first[i]] : -1;
    }
    __syncthreads();
    for (int x = tid; x < 2 * n; x += blockDim.x) {
        orig[x] = (x < n) ? k1 + x : ((first[x - n] == x - n) ? tgt[x - n] : -1);
        last[x] = -1;
        cnt[x] = 0;
    }
    for (int i = tid; i < n; i += blockDim.x) {
        const int t = tgt[i];
        const int sl = (t >= k1 && t < k2) ? (t - k1) : (n + first[i]);
        slot[i] = sl;
        if (sl < i) s_back = 1;
    }
    __syncthreads();
    if (s_back) {
        // in-order replay on orig (becomes the contents)
        if (tid == 0) {
            for (int i = 0; i < n; i++) {
                const int sl = slot[i];
                if (sl != i) {
                    const int x = orig[i];
                    orig[i] = orig[sl];
                    orig[sl] = x;
                }
            }
        }
        __syncthreads();
        for (int x = tid; x < 2 * n; x += blockDim.x) {
            const int row = (x < n) ? k1 + x : ((first[x - n] == x - n) ? tgt[x - n] : -1);
            pairs[2 * x] = row;
            pairs[2 * x + 1] = (row >= 0) ? orig[x] : -1;
        }
        if (tid == 0) *count = 2 * n;
        return;
    }
    for (int i = tid; i < n; i += blockDim.x) {
        const int sl = slot[i];
        if (sl != i) {
            atomicMax(&last[sl], i);
            atomicAdd(&cnt[sl], 1);
        }
    }
    __syncthreads();
    // the previous step with the same slot: a scan only where the slot is
    // shared (a row chosen as pivot twice)
    for (int i = tid; i < n; i += blockDim.x) {
        const int sl = slot[i];
        int l = -1;
        if (sl != i && cnt[sl] > 1)
            for (int j = i - 1; j >= 0; j--)
                if (slot[j] == sl) {
                    l = j;
                    break;
                }
        prev[i] = l;
    }
    __syncthreads();
    // n <= 512 = blockDim: one step per thread
    const bool mine = tid < n;
    int p = 0;
    if (mine) {
        p = (last[tid] >= 0) ? last[tid] : tid;
        jump[tid] = p;
    }
    __syncthreads();
    for (int span = 1; span < n; span *= 2) {
        const int q = mine ? jump[p] : 0;
        __syncthreads();
        if (mine) {
            p = q;
            jump[tid] = p;
        }
        __syncthreads();
    }
    if (mine) val[tid] = orig[p];
    __syncthreads();
    for (int x = tid; x < 2 * n; x += blockDim.x) {
        const int row = orig[x];
        int f;
        if (x < n) {
            const int sl = slot[x];
            const int l = prev[x];
            f = (sl == x) ? val[x] : (l >= 0 ? val[l] : orig[sl]);
        } else {
            const int l = last[x];
            f = (l >= 0) ? val[l] : row;
        }
        pairs[2 * x] = row;
        pairs[2 * x + 1] = (row >= 0) ? f : -1;
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
__global__ void reset_kstamps_kernel(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[2 * i] = ~0ull;
        p[2 * i + 1] = 0ull;
    }
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
void reset_kstamps(unsigned long long *p, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(reset_kstamps_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
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
