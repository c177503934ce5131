// mixed.hip -- mixed-precision solve: factor a copy of A in FP32 storage
// (trailing updates in FP32, BF16 or FP8 on MFMA), then refine the solution
// until it has FP64 backward error.  Only the residual r = b - A x, the
// solution update and the convergence test run in FP64; the corrections are
// solves with the low-precision factors, either directly (classical iterative
// refinement) or as the preconditioner of a flexible GMRES on A d = r
// (GMRES-IR, for ill-conditioned A).
//
// The contract follows LAPACK's DSGESV (the mixed-precision counterpart of
// the reference's PDGESV, SRC/pdgesv.f): A is left unchanged when refinement
// succeeds; otherwise A is factored in FP64 in place, X is solved from it and
// ITER is negative:
//   -2   converting A to FP32 overflowed
//   -3   the low-precision factorization hit an exactly zero pivot
//   -(max_iter+1)  refinement did not converge (max_iter steps, or a stall
//        above the threshold; DSGESV's -31 for its ITERMAX = 30)
// Convergence test (DSGESV's): ||r||_inf <= ||x||_inf ||A||_inf sqrt(N) eps
// per column, eps = 2^-53.  A step that does not at least halve the
// normwise backward error counts as a stall (the stopping rule of PDGERFS,
// SRC/pdgerfs.f:556-561); a stalled result is accepted when its scaled
// residual ||r|| / (||A|| ||x|| N eps) is below the testers' threshold 16
// (TESTING/traditional/LIN/pdlaschk.f:296), unless the handle's
// ir_accept_stall is 0.  MixedInfo::stop_rule records which rule ended it.
#include "sla_common.h"
#include "sla_kernels.h"
#include "sla_handle.h"
#include "sla_mixed.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace sla {

namespace {

constexpr double EPS64 = 1.1102230246251565e-16;  // 2^-53

// ---- FP64 residual: R = B - A X (n x nrhs), A n x n --------------------------
// Stage 1: each workgroup takes 512 rows (2 per thread, one 16-byte load per
// column) and a slice of KC columns and writes its partial products.
constexpr int RES_ROWS = 512, RES_KC = 512;

template <int NR>
__global__ __launch_bounds__(256) void gemv_partial_kernel(int n, int k, int nr, bool vec,
                                                           const double *__restrict__ A, int lda,
                                                           const double *__restrict__ X, int ldx,
                                                           double *__restrict__ part)
{
    __shared__ double xs[RES_KC * NR];
    const int r0 = blockIdx.x * RES_ROWS, c0 = blockIdx.y * RES_KC;
    const int kc = min(RES_KC, k - c0);
    for (int idx = threadIdx.x; idx < kc * NR; idx += 256) {
        const int kk = idx % kc, j = idx / kc;
        xs[kk * NR + j] = j < nr ? X[c0 + kk + (size_t)j * ldx] : 0.0;
    }
    __syncthreads();
    const int row = r0 + 2 * threadIdx.x;
    double acc0[NR], acc1[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) acc0[j] = acc1[j] = 0.0;
    if (vec && row + 1 < n) {
        const double *Ap = A + row + (size_t)c0 * lda;
        int kk = 0;
        for (; kk + 4 <= kc; kk += 4) {
            double2 a[4];
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] = *(const double2 *)(Ap + (size_t)(kk + u) * lda);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int j = 0; j < NR; j++) {
                    acc0[j] = fma(a[u].x, xs[(kk + u) * NR + j], acc0[j]);
                    acc1[j] = fma(a[u].y, xs[(kk + u) * NR + j], acc1[j]);
                }
        }
        for (; kk < kc; kk++) {
            const double2 a = *(const double2 *)(Ap + (size_t)kk * lda);
#pragma unroll
            for (int j = 0; j < NR; j++) {
                acc0[j] = fma(a.x, xs[kk * NR + j], acc0[j]);
                acc1[j] = fma(a.y, xs[kk * NR + j], acc1[j]);
            }
        }
    } else {
        // odd leading dimension or the last odd row: one row at a time
        const double *Ap = A + row + (size_t)c0 * lda;
        if (row < n)
            for (int kk = 0; kk < kc; kk++) {
                const double a = Ap[(size_t)kk * lda];
#pragma unroll
                for (int j = 0; j < NR; j++) acc0[j] = fma(a, xs[kk * NR + j], acc0[j]);
            }
        if (row + 1 < n)
            for (int kk = 0; kk < kc; kk++) {
                const double a = Ap[1 + (size_t)kk * lda];
#pragma unroll
                for (int j = 0; j < NR; j++) acc1[j] = fma(a, xs[kk * NR + j], acc1[j]);
            }
    }
    // part layout: [slice][j][n]
    double *P = part + (size_t)blockIdx.y * NR * n;
#pragma unroll
    for (int j = 0; j < NR; j++) {
        if (row < n) P[(size_t)j * n + row] = acc0[j];
        if (row + 1 < n) P[(size_t)j * n + row + 1] = acc1[j];
    }
}

// Stage 2: R = alpha * B - sum of the slices (fixed order, so the residual is
// reproducible bit for bit); column max |R| into rmax (bit-pattern max).
template <int NR>
__global__ __launch_bounds__(256) void gemv_reduce_kernel(int n, int nr, int slices, double alpha,
                                                          const double *__restrict__ B, int ldb,
                                                          const double *__restrict__ part,
                                                          double *__restrict__ R, int ldr,
                                                          unsigned long long *rmax)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    for (int j = 0; j < nr; j++) {
        double v = 0.0;
        if (row < n) {
            double s = 0.0;
            for (int c = 0; c < slices; c++) s += part[((size_t)c * NR + j) * n + row];
            v = (alpha != 0.0 ? alpha * B[row + (size_t)j * ldb] : 0.0) - s;
            R[row + (size_t)j * ldr] = v;
        }
        double a = fabs(v);
        for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o));
        if (rmax && (threadIdx.x & 63) == 0)
            atomicMax(rmax + j, (unsigned long long)__double_as_longlong(a));
    }
}

// ---- conversions and updates -----------------------------------------------
// W(i, j) = (float) R(perm[i], j): the row interchanges of the factorization
// and the narrowing in one pass (PDLAPIV's gather, SRC/pdlapiv.f:242-244).
__global__ void gather_narrow_kernel(int n, int nrhs, const double *__restrict__ R, int ldr,
                                     const int *__restrict__ perm, float *__restrict__ W, int ldw)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    W[i + (size_t)j * ldw] = (float)R[perm[i] + (size_t)j * ldr];
}

// X(:, j) += D(:, j) (FP32 correction), column max |X| into xmax.
__global__ __launch_bounds__(256) void update_kernel(int n, int nrhs, const float *__restrict__ D,
                                                     int ldd, double *__restrict__ X, int ldx,
                                                     unsigned long long *xmax)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    for (int j = 0; j < nrhs; j++) {
        double a = 0.0;
        if (row < n) {
            const double x = X[row + (size_t)j * ldx] + (double)D[row + (size_t)j * ldd];
            X[row + (size_t)j * ldx] = x;
            a = fabs(x);
        }
        for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o));
        if ((threadIdx.x & 63) == 0) atomicMax(xmax + j, (unsigned long long)__double_as_longlong(a));
    }
}

// FP32 copy of A plus the row sums of |A| (||A||_inf is their max).
constexpr int NARROW_COLS = 256;

__global__ __launch_bounds__(256) void narrow_rowsum_kernel(int n, const double *__restrict__ A,
                                                            int lda, float *__restrict__ S, int lds,
                                                            double *__restrict__ rowsum,
                                                            int *overflow)
{
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * NARROW_COLS, c1 = min(n, c0 + NARROW_COLS);
    if (row >= n) return;
    double acc = 0.0;
    bool bad = false;
    for (int c = c0; c < c1; c++) {
        const double a = A[row + (size_t)c * lda];
        bad |= !(fabs(a) <= 3.4028234663852886e38);
        acc += fabs(a);
        S[row + (size_t)c * lds] = (float)a;
    }
    atomicAdd(rowsum + row, acc);
    if (bad) *overflow = 1;
}

__global__ __launch_bounds__(256) void max_kernel(int n, const double *__restrict__ v,
                                                  unsigned long long *out)
{
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a = fmax(a, v[i]);
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(a));
}

__global__ void fill_f64_kernel(double *p, size_t n, double v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- small dense kernels of the flexible GMRES ------------------------------
// h[j] = V(:, j)^T w for j < nj: one workgroup per column, fixed-order tree.
__global__ __launch_bounds__(256) void multidot_kernel(int n, int nj, const double *__restrict__ V,
                                                       int ldv, const double *__restrict__ w,
                                                       double *__restrict__ h)
{
    __shared__ double red[4];
    const int j = blockIdx.x;
    const double *v = V + (size_t)j * ldv;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s = fma(v[i], w[i], s);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) h[j] = (red[0] + red[1]) + (red[2] + red[3]);
}

// w -= V(:, 0:nj) h
__global__ __launch_bounds__(256) void gemv_sub_kernel(int n, int nj, const double *__restrict__ V,
                                                       int ldv, const double *__restrict__ h,
                                                       double *__restrict__ w)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < nj; j++) s = fma(V[i + (size_t)j * ldv], h[j], s);
    w[i] -= s;
}

// y = alpha * x (FP64) or x += Z y
__global__ void scale_copy_kernel(int n, double alpha, const double *__restrict__ x,
                                  double *__restrict__ y)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = alpha * x[i];
}

__global__ __launch_bounds__(256) void combine_kernel(int n, int nj, const double *__restrict__ Z,
                                                      int ldz, const double *__restrict__ y,
                                                      double *__restrict__ x,
                                                      unsigned long long *xmax)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    double a = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int j = 0; j < nj; j++) s = fma(Z[i + (size_t)j * ldz], y[j], s);
        const double v = x[i] + s;
        x[i] = v;
        a = fabs(v);
    }
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o));
    if ((threadIdx.x & 63) == 0) atomicMax(xmax, (unsigned long long)__double_as_longlong(a));
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }
size_t round_up(size_t n, size_t q) { return (n + q - 1) / q * q; }

// ---------------------------------------------------------------------------
struct Mixed {
    Handle &h;
    int n, nrhs;
    double *A;
    int lda;
    const double *B;
    int ldb;
    double *X;
    int ldx;
    float *S = nullptr;  // FP32 factors
    int lds = 0;
    const double *dinv = nullptr;  // inverses of their diagonal blocks (null: substitution)
    int *ipiv_s = nullptr;
    double anorm = 0.0;

    // vectors: R (n x nrhs FP64), W (n x max(nrhs,1) FP32), partial sums
    double *R = nullptr;
    float *W = nullptr;
    double *part = nullptr;
    unsigned long long *norms = nullptr;  // [0, nrhs): |r|max, [16, 16+nrhs): |x|max
    int slices = 0;

    hipStream_t st() const { return h.stream; }

    void alloc()
    {
        const size_t nn = (size_t)n;
        // R, then W, then the norm words; every part starts on a 256-byte
        // boundary (the norms take 64-bit atomics, which must be naturally
        // aligned: 12 n nrhs bytes alone is not a multiple of 8 for odd n nrhs)
        const size_t offw = round_up(nn * nrhs * 8, 256);
        const size_t offn = offw + round_up(nn * nrhs * 4, 256);
        char *v = static_cast<char *>(h.buffer(BUF_VEC, offn + 4096));
        R = reinterpret_cast<double *>(v);
        W = reinterpret_cast<float *>(v + offw);
        norms = reinterpret_cast<unsigned long long *>(v + offn);
        slices = ceil_div(n, RES_KC);
        const int nrp = nrhs <= 1 ? 1 : nrhs <= 4 ? 4 : 16;
        part = static_cast<double *>(h.buffer(BUF_PARTIAL, (size_t)slices * nrp * nn * 8 + 256));
    }

    // R(:, 0:nr) = alpha * Bc - A Xc, column max |R| into norms[0:nr].
    void residual(const double *Bc, int ldbc, double alpha, const double *Xc, int ldxc, double *Rc,
                  int ldrc, int nr, unsigned long long *rmax)
    {
        const dim3 g1(ceil_div(n, RES_ROWS), slices), g2(blocks_for(n));
        if (nr < 1 || nr > 16) throw std::logic_error("residual: 1..16 columns");
        const bool vec = (lda % 2 == 0) && (reinterpret_cast<uintptr_t>(A) % 16 == 0);
        h.timing_begin(CAT_OTHER);
        if (nr == 1) {
            hipLaunchKernelGGL(gemv_partial_kernel<1>, g1, dim3(256), 0, st(), n, n, nr, vec, A, lda,
                               Xc, ldxc, part);
            hipLaunchKernelGGL(gemv_reduce_kernel<1>, g2, dim3(256), 0, st(), n, nr, slices, alpha,
                               Bc, ldbc, part, Rc, ldrc, rmax);
        } else if (nr <= 4) {
            hipLaunchKernelGGL(gemv_partial_kernel<4>, g1, dim3(256), 0, st(), n, n, nr, vec, A, lda,
                               Xc, ldxc, part);
            hipLaunchKernelGGL(gemv_reduce_kernel<4>, g2, dim3(256), 0, st(), n, nr, slices, alpha,
                               Bc, ldbc, part, Rc, ldrc, rmax);
        } else {
            hipLaunchKernelGGL(gemv_partial_kernel<16>, g1, dim3(256), 0, st(), n, n, nr, vec, A, lda,
                               Xc, ldxc, part);
            hipLaunchKernelGGL(gemv_reduce_kernel<16>, g2, dim3(256), 0, st(), n, nr, slices,
                               alpha, Bc, ldbc, part, Rc, ldrc, rmax);
        }
        SLA_LAUNCH_CHECK();
        h.timing_end(8.0 * (double)n * n);
    }

    // W = U^{-1} L^{-1} P R in FP32 (nr <= 16 columns), low-precision factors.
    void lowp_solve(const double *Rc, int ldrc, int nr)
    {
        h.timing_begin(CAT_LASWP);
        hipLaunchKernelGGL(gather_narrow_kernel, dim3(blocks_for((size_t)n * nr)), dim3(256), 0,
                           st(), n, nr, Rc, ldrc, h.perm, W, n);
        SLA_LAUNCH_CHECK();
        h.timing_end(12.0 * n * nr);
        TrsvWork tw = h.trsv_work(n);
        for (const bool lower : {true, false}) {
            h.timing_begin(CAT_TRSV);
            if (dinv) trsv_blockinv<float>(lower, n, nr, S, lds, dinv, W, n, tw, st());
            else trsv_persistent<float, float>(lower ? TrsvOp::L : TrsvOp::U, n, nr, S, lds, W, n, tw, st());
            h.timing_end(2.0 * (double)n * n);
        }
        h.trsv_epoch = tw.epoch;
    }

    // Z = U^{-1} L^{-1} P V (one column, FP64 arithmetic on the FP32
    // factors: GMRES-IR's preconditioner, applied in the working precision).
    void precond64(const double *v, double *z)
    {
        h.timing_begin(CAT_LASWP);
        gather_rows<double>(v, n, z, n, h.perm, n, 1, st());
        h.timing_end(16.0 * n);
        TrsvWork tw = h.trsv_work(n);
        for (const bool lower : {true, false}) {
            h.timing_begin(CAT_TRSV);
            if (dinv) trsv_blockinv<double>(lower, n, 1, S, lds, dinv, z, n, tw, st());
            else trsv_persistent<float, double>(lower ? TrsvOp::L : TrsvOp::U, n, 1, S, lds, z, n, tw, st());
            h.timing_end(2.0 * (double)n * n);
        }
        h.trsv_epoch = tw.epoch;
    }

    void check_flags()
    {
        int flags[2];
        SLA_HIP(hipMemcpyAsync(flags, h.ctl, sizeof flags, hipMemcpyDeviceToHost, st()));
        SLA_HIP(hipStreamSynchronize(st()));
        if (flags[1]) throw std::runtime_error("mixed solve: a dependency wait timed out");
    }

    std::vector<double> read_norms(int off, int nr)
    {
        std::vector<unsigned long long> bits(nr);
        SLA_HIP(hipMemcpyAsync(bits.data(), norms + off, nr * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, st()));
        SLA_HIP(hipStreamSynchronize(st()));
        std::vector<double> v(nr);
        for (int j = 0; j < nr; j++) memcpy(&v[j], &bits[j], 8);
        return v;
    }

    void zero_norms(int off, int nr)
    {
        hipLaunchKernelGGL(fill_f64_kernel, dim3(1), dim3(64), 0, st(),
                           reinterpret_cast<double *>(norms + off), (size_t)nr, 0.0);
        SLA_LAUNCH_CHECK();
    }

    // One correction of all columns: X += solve(R) (classical refinement).
    void correct_classical()
    {
        for (int j = 0; j < nrhs; j += 16) {
            const int nr = std::min(16, nrhs - j);
            lowp_solve(R + (size_t)j * n, n, nr);
            hipLaunchKernelGGL(update_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n, nr, W, n,
                               X + (size_t)j * ldx, ldx, norms + 16 + j);
            SLA_LAUNCH_CHECK();
        }
    }

    // Flexible GMRES on A d = r for column j, right-preconditioned by the
    // low-precision factors, started from d = 0; X(:, j) += d.  Returns the
    // number of Arnoldi steps taken.
    int correct_gmres(int j, int mmax, double tol)
    {
        static const bool trace = getenv("SLA_TRACE_GMRES") != nullptr;
        const size_t nn = (size_t)n;
        double *V = static_cast<double *>(h.buffer(BUF_KRYLOV, nn * (2 * mmax + 1) * 8 + 8 * 4096));
        double *Z = V + nn * (mmax + 1);
        double *hd = Z + nn * mmax;  // device scratch for dot products (<= mmax + 1)
        const double *r = R + (size_t)j * n;
        // beta = ||r||_2
        hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(256), 0, st(), n, 1, r, n, r, hd);
        SLA_LAUNCH_CHECK();
        double beta = 0.0;
        SLA_HIP(hipMemcpyAsync(&beta, hd, 8, hipMemcpyDeviceToHost, st()));
        SLA_HIP(hipStreamSynchronize(st()));
        beta = std::sqrt(beta);
        if (!(beta > 0.0)) return 0;
        hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n,
                           1.0 / beta, r, V);
        SLA_LAUNCH_CHECK();
        std::vector<double> H((size_t)(mmax + 1) * mmax, 0.0), cs(mmax), sn(mmax), g(mmax + 1, 0.0);
        g[0] = beta;
        int k = 0;
        std::vector<double> hcol(mmax + 1), hcol2(mmax + 1);
        for (; k < mmax; k++) {
            double *vk = V + nn * k, *zk = Z + nn * k, *w = V + nn * (k + 1);
            // z_k = M^{-1} v_k
            precond64(vk, zk);
            // w = A z_k (FP64)
            residual(nullptr, n, 0.0, zk, n, w, n, 1, nullptr);
            hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n, -1.0,
                               w, w);
            // classical Gram-Schmidt, twice
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL(multidot_kernel, dim3(k + 1), dim3(256), 0, st(), n, k + 1, V,
                                   n, w, hd);
                hipLaunchKernelGGL(gemv_sub_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n,
                                   k + 1, V, n, hd, w);
                SLA_HIP(hipMemcpyAsync(pass ? hcol2.data() : hcol.data(), hd, (k + 1) * 8,
                                       hipMemcpyDeviceToHost, st()));
            }
            hipLaunchKernelGGL(multidot_kernel, dim3(1), dim3(256), 0, st(), n, 1, w, n, w, hd);
            SLA_LAUNCH_CHECK();
            double hk1 = 0.0;
            SLA_HIP(hipMemcpyAsync(&hk1, hd, 8, hipMemcpyDeviceToHost, st()));
            SLA_HIP(hipStreamSynchronize(st()));
            hk1 = std::sqrt(hk1);
            for (int i = 0; i <= k; i++) H[(size_t)k * (mmax + 1) + i] = hcol[i] + hcol2[i];
            H[(size_t)k * (mmax + 1) + k + 1] = hk1;
            if (hk1 > 0.0) {
                hipLaunchKernelGGL(scale_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n,
                                   1.0 / hk1, w, w);
                SLA_LAUNCH_CHECK();
            }
            // Givens rotations on the new column
            double *hc = &H[(size_t)k * (mmax + 1)];
            for (int i = 0; i < k; i++) {
                const double t = cs[i] * hc[i] + sn[i] * hc[i + 1];
                hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
                hc[i] = t;
            }
            const double den = std::hypot(hc[k], hc[k + 1]);
            cs[k] = den > 0.0 ? hc[k] / den : 1.0;
            sn[k] = den > 0.0 ? hc[k + 1] / den : 0.0;
            hc[k] = den;
            hc[k + 1] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            if (trace) fprintf(stderr, "gmres step %d: |g|/beta = %.3e  h(k+1,k) = %.3e\n", k + 1,
                               std::fabs(g[k + 1]) / beta, hk1);
            if (std::fabs(g[k + 1]) <= tol * beta || hk1 == 0.0) {
                k++;
                break;
            }
        }
        // y = H(0:k, 0:k)^{-1} g, back substitution
        std::vector<double> y(k, 0.0);
        for (int i = k - 1; i >= 0; i--) {
            double s = g[i];
            for (int c = i + 1; c < k; c++) s -= H[(size_t)c * (mmax + 1) + i] * y[c];
            y[i] = s / H[(size_t)i * (mmax + 1) + i];
        }
        SLA_HIP(hipMemcpyAsync(hd, y.data(), k * 8, hipMemcpyHostToDevice, st()));
        hipLaunchKernelGGL(combine_kernel, dim3(blocks_for(n)), dim3(256), 0, st(), n, k, Z, n, hd,
                           X + (size_t)j * ldx, norms + 16 + j);
        SLA_LAUNCH_CHECK();
        SLA_HIP(hipStreamSynchronize(st()));
        return k;
    }
};

}  // namespace

void dresidual_device(Handle &h, int n, int nrhs, const double *dA, int lda, const double *dX,
                      int ldx, const double *dB, int ldb, double *dR, int ldr, double *rnorm)
{
    if (n <= 0 || nrhs <= 0) return;
    Mixed mx{h, n, nrhs, const_cast<double *>(dA), lda, dB, ldb, nullptr, 0};
    mx.alloc();
    for (int j = 0; j < nrhs; j += 16) {
        const int nr = std::min(16, nrhs - j);
        mx.zero_norms(0, nr);
        mx.residual(dB + (size_t)j * ldb, ldb, 1.0, dX + (size_t)j * ldx, ldx,
                    dR + (size_t)j * ldr, ldr, nr, mx.norms);
        const std::vector<double> v = mx.read_norms(0, nr);
        for (int c = 0; c < nr; c++) rnorm[j + c] = v[c];
    }
}

void sgetrs_device(Handle &h, int n, int nrhs, const float *dA, int lda, const int *dipiv,
                   float *dB, int ldb)
{
    if (n <= 0 || nrhs <= 0) return;
    // flags of this call only: an earlier call's out-of-range IPIV or timeout
    // must neither fail this one nor survive it
    fill_int(h.bad_pivot, 1, 0, h.stream);
    fill_int(h.abort_flag, 1, 0, h.stream);
    if (!(h.perm_key_a == dA && h.perm_key_ipiv == dipiv && h.perm_m == n))
        perm_from_ipiv(h, n, dipiv);
    float *W = static_cast<float *>(h.buffer(BUF_VEC, (size_t)n * 16 * 4 + 256));
    double *inv = nullptr;
    if (h.trsv_blockinv) {
        inv = static_cast<double *>(h.buffer(BUF_DINV, trsv_inverse_doubles(n) * 8));
        trsv_block_inverses(n, dA, lda, inv, h.stream);
    }
    TrsvWork tw = h.trsv_work(n);
    for (int j = 0; j < nrhs; j += 16) {
        const int nr = std::min(16, nrhs - j);
        float *Bj = dB + (size_t)j * ldb;
        gather_rows<float>(Bj, ldb, W, n, h.perm, n, nr, h.stream);
        if (inv) {
            trsv_blockinv<float>(true, n, nr, dA, lda, inv, W, n, tw, h.stream);
            trsv_blockinv<float>(false, n, nr, dA, lda, inv, W, n, tw, h.stream);
        } else {
            trsv_persistent<float, float>(TrsvOp::L, n, nr, dA, lda, W, n, tw, h.stream);
            trsv_persistent<float, float>(TrsvOp::U, n, nr, dA, lda, W, n, tw, h.stream);
        }
        SLA_HIP(hipMemcpy2DAsync(Bj, (size_t)ldb * 4, W, (size_t)n * 4, (size_t)n * 4, nr,
                                 hipMemcpyDeviceToDevice, h.stream));
    }
    h.trsv_epoch = tw.epoch;
    // an IPIV row outside 1..N leaves the permutation incomplete: an error,
    // never a wrong X (the same check as dgetrs)
    check_solve_flags(h, "sgetrs");
}

int dxgesv_device(Handle &h, int prec, int refine, int n, int nrhs, double *dA, int lda,
                  int *dipiv, const double *dB, int ldb, double *dX, int ldx, MixedInfo &info)
{
    info = MixedInfo{};
    info.precision = prec;
    info.refinement = refine;
    if (n <= 0 || nrhs <= 0) return 0;
    if (nrhs > 16) throw std::invalid_argument("dxgesv: at most 16 right-hand sides per call");
    const double cte = std::sqrt((double)n) * EPS64;

    auto fallback = [&](int iter) {
        if (iter != -1 && !h.ir_fallback) {
            // the caller measures the attempt alone: report the failure
            info.stop_rule = STOP_FAILED;
            info.iter = iter;
            h.timing_collect();
            return 0;
        }
        // FP64 factorization in place and solve of X = A^{-1} B; A then holds
        // the factors, so the errors of this X are not measured (NaN)
        info.fell_back = 1;
        info.stop_rule = STOP_FALLBACK;
        info.iter = iter;
        info.bwd_err = info.sresid = std::nan("");
        SLA_HIP(hipMemcpy2DAsync(dX, (size_t)ldx * 8, dB, (size_t)ldb * 8, (size_t)n * 8, nrhs,
                                 hipMemcpyDeviceToDevice, h.stream));
        const int zi = dgetrf_device(h, n, n, dA, lda, dipiv);
        if (zi > 0) return zi;
        dgetrs_device(h, n, nrhs, dA, lda, dipiv, dX, ldx);
        h.timing_collect();
        return 0;
    };

    if (prec == PREC_FP64) return fallback(-1);

    Mixed mx{h, n, nrhs, dA, lda, dB, ldb, dX, ldx};
    mx.lds = ceil_div(n, 64) * 64;
    mx.S = static_cast<float *>(h.buffer(BUF_SA, (size_t)mx.lds * n * 4));
    mx.ipiv_s = static_cast<int *>(h.buffer(BUF_IPIV_S, (size_t)n * 4));
    mx.alloc();

    // FP32 copy of A and ||A||_inf in one pass over A
    {
        double *rowsum = mx.part;  // n doubles, before the partial sums are needed
        hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks_for(n)), dim3(256), 0, h.stream, rowsum,
                           (size_t)n, 0.0);
        mx.zero_norms(0, 32);
        fill_int(h.bad_pivot, 1, 0, h.stream);
        h.timing_begin(CAT_OTHER);
        hipLaunchKernelGGL(narrow_rowsum_kernel, dim3(blocks_for(n), ceil_div(n, NARROW_COLS)),
                           dim3(256),
                           0, h.stream, n, dA, lda, mx.S, mx.lds, rowsum, h.bad_pivot);
        hipLaunchKernelGGL(max_kernel, dim3(64), dim3(256), 0, h.stream, n, rowsum, mx.norms);
        SLA_LAUNCH_CHECK();
        h.timing_end(12.0 * (double)n * n);
        int ovf = 0;
        SLA_HIP(hipMemcpyAsync(&ovf, h.bad_pivot, 4, hipMemcpyDeviceToHost, h.stream));
        mx.anorm = mx.read_norms(0, 1)[0];
        info.anorm = mx.anorm;
        if (ovf) return fallback(-2);
    }
    const int schur = schur_of_prec(prec);
    if (sgetrf_device(h, n, n, mx.S, mx.lds, mx.ipiv_s, schur) > 0) return fallback(-3);
    if (h.trsv_blockinv) {
        double *inv = static_cast<double *>(h.buffer(BUF_DINV, trsv_inverse_doubles(n) * 8));
        h.timing_begin(CAT_TRSV);
        trsv_block_inverses(n, mx.S, mx.lds, inv, h.stream);
        h.timing_end(0.0);
        mx.dinv = inv;
    }

    // x0 = solve(b): X = 0, R = B
    hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks_for((size_t)n * nrhs)), dim3(256), 0, h.stream,
                       mx.R, (size_t)n * nrhs, 0.0);
    for (int j = 0; j < nrhs; j++)
        hipLaunchKernelGGL(fill_f64_kernel, dim3(blocks_for(n)), dim3(256), 0, h.stream,
                           dX + (size_t)j * ldx, (size_t)n, 0.0);
    SLA_HIP(hipMemcpy2DAsync(mx.R, (size_t)n * 8, dB, (size_t)ldb * 8, (size_t)n * 8, nrhs,
                             hipMemcpyDeviceToDevice, h.stream));
    mx.zero_norms(16, nrhs);
    mx.correct_classical();

    std::vector<double> last_bwd(nrhs, HUGE_VAL);
    bool gmres = refine == REFINE_GMRES;
    for (int it = 0;; it++) {
        mx.zero_norms(0, nrhs);
        mx.residual(dB, ldb, 1.0, dX, ldx, mx.R, n, nrhs, mx.norms);
        mx.check_flags();
        const std::vector<double> rn = mx.read_norms(0, nrhs), xn = mx.read_norms(16, nrhs);
        bool converged = true, stalled = false, acceptable = true;
        double worst = 0.0;
        for (int j = 0; j < nrhs; j++) {
            const double denom = xn[j] * mx.anorm;
            const double bwd = denom > 0.0 ? rn[j] / denom : (rn[j] > 0.0 ? HUGE_VAL : 0.0);
            worst = std::max(worst, bwd);
            if (!(rn[j] <= xn[j] * mx.anorm * cte)) converged = false;
            if (!(bwd <= 0.5 * last_bwd[j])) stalled = true;
            if (!(bwd / (n * EPS64) < 16.0)) acceptable = false;
            last_bwd[j] = bwd;
        }
        info.bwd_err = worst;
        info.sresid = worst / (n * EPS64);
        info.outer = it;
        if (it < 16) info.history[it] = worst;
        if (converged || (stalled && acceptable && h.ir_accept_stall)) {
            info.iter = it;
            info.stop_rule = converged ? STOP_CONVERGED : STOP_STALL_ACCEPTED;
            h.timing_collect();
            return 0;
        }
        if (it >= h.ir_max_iter) break;
        if (stalled) {
            // classical refinement stalled above the threshold: continue with
            // GMRES-IR if allowed, else give up
            if (gmres || !h.ir_gmres_on_stall) break;
            gmres = true;
        }
        mx.zero_norms(16, nrhs);
        if (gmres) {
            info.refinement = REFINE_GMRES;
            for (int j = 0; j < nrhs; j++) info.inner += mx.correct_gmres(j, h.gmres_m, h.gmres_tol);
        } else {
            mx.correct_classical();
        }
    }
    return fallback(-(h.ir_max_iter + 1));
}

}  // namespace sla
