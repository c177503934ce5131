// trsm.hip -- triangular solves of the LU path.
//
// trsm_llnu: the block row of U in PDGETRF, U12 <- L11^{-1} A12
//   (SRC/pdgetrf.f:280-281, PDTRSM('L','L','N','U'); on a 1 x 1 grid
//   PBLAS/SRC/PTOOLS/PB_CptrsmAB.c:359-412 splits it into 32-row dtrsm_ +
//   dgemm_ steps).  Each workgroup owns a strip of 16 columns of A12, keeps
//   it in LDS, and walks L11 in 32-column blocks staged through LDS: solve the
//   32 x 32 diagonal block, then update the rows below with one fma chain.
//
// trsv_persistent: the two solves of PDGETRS (SRC/pdgetrs.f:260-266 for
//   TRANS = 'N', :273-284 for 'T') and of PDLATRS inside PDGECON
//   (SRC/pdgecon.f:342-370) for a few right-hand sides.  Those are
//   latency-bound, so one launch covers the whole triangle: block row b
//   (64 rows) is owned by the workgroup that takes ticket b; it accumulates
//   -T[b, c] X[c] for every finished block c (X[c] read from a mailbox whose
//   words carry the launch's epoch, bounded wait), then solves its 64 x 64
//   diagonal block in registers of one wave (lane = row, values broadcast by
//   lane reads) and publishes X[b].
//   Tickets are taken in start order, so a workgroup only ever waits on
//   workgroups that are already running.  The transposed solves read the
//   off-diagonal block by columns (lane = column of the stored triangle,
//   16 consecutive elements of that column per wave).
#include "sla_common.h"
#include "sla_kernels.h"

#include <type_traits>

namespace sla {

namespace {

constexpr int TS_CW = 16, TS_KB = 32, TS_MAXM = 256;

template <typename T>
__global__ __launch_bounds__(256) void trsm_llnu_strip_kernel(int m, int n, const T *__restrict__ L,
                                                              int ldl, T *__restrict__ B, int ldb)
{
    __shared__ T Bs[TS_MAXM][TS_CW + 1];
    __shared__ T Ls[TS_MAXM][TS_KB + 1];
    const int c0 = blockIdx.x * TS_CW, tid = threadIdx.x;
    const int c = tid % TS_CW, rg = tid / TS_CW;  // 16 row groups
    for (int idx = tid; idx < m * TS_CW; idx += 256) {
        const int r = idx % m, cc = idx / m;
        Bs[r][cc] = (c0 + cc < n) ? B[r + (size_t)(c0 + cc) * ldb] : T(0);
    }
    for (int kb = 0; kb < m; kb += TS_KB) {
        const int kbb = min(TS_KB, m - kb), rows = m - kb;
        __syncthreads();  // previous block's readers of Ls are done
        for (int idx = tid; idx < rows * kbb; idx += 256) {
            const int r = idx % rows, k = idx / rows;
            Ls[r][k] = L[(size_t)(kb + r) + (size_t)(kb + k) * ldl];
        }
        __syncthreads();
        for (int i = 0; i < kbb; i++) {
            const T x = Bs[kb + i][c];
            for (int r = i + 1 + rg; r < kbb; r += 16) Bs[kb + r][c] = fma(-Ls[r][i], x, Bs[kb + r][c]);
            __syncthreads();
        }
        for (int r = kbb + rg; r < rows; r += 16) {
            T acc = Bs[kb + r][c];
            for (int k = 0; k < kbb; k++) acc = fma(-Ls[r][k], Bs[kb + k][c], acc);
            Bs[kb + r][c] = acc;
        }
    }
    __syncthreads();
    for (int idx = tid; idx < m * TS_CW; idx += 256) {
        const int r = idx % m, cc = idx / m;
        if (c0 + cc < n) B[r + (size_t)(c0 + cc) * ldb] = Bs[r][cc];
    }
}

// ---------------------------------------------------------------------------
constexpr int TV_BS = 64, TV_NR = 16, TV_PART = TV_BS / 4;

// The stored matrix holds L (unit lower, below the diagonal) and U (upper).
//   !TRANS,  FORWARD: X = L^{-1} X        !TRANS, !FORWARD: X = U^{-1} X
//    TRANS,  FORWARD: X = U^{-T} X         TRANS, !FORWARD: X = L^{-T} X
// A non-unit diagonal is applied as a product with its reciprocals, computed
// before any wait so they stay off the dependency chain.  X is n x nrhs
// (nrhs <= NR <= 16) with leading dimension ldx; the matrix is n x n.
// TA: storage type of the triangle; T: type of X and of the arithmetic (a
// FP32 factorization may be applied in FP64, as GMRES-IR's preconditioner).
//
// Hand-off between blocks: a solved block publishes its values in a mailbox
// of 64-bit words, each holding 32 bits of a value and the launch's epoch in
// the upper half (a double takes two words), written with agent-coherent
// stores.  A waiting workgroup polls the words themselves (wave 0, lane = row)
// until every one carries the epoch -- one memory round trip from producer to
// consumer, no separate flag and no store drain.  Stale words of earlier
// launches hold older epochs; the mailbox is zeroed when allocated.  The next
// off-diagonal block of the triangle is loaded into registers before the
// wait, so the matrix read stays off the chain.
template <typename T>
__device__ __forceinline__ uint32_t tv_half(T v, int w)
{
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        return w ? (uint32_t)(u >> 32) : (uint32_t)u;
    } else {
        return __builtin_bit_cast(uint32_t, v);
    }
}

template <typename T>
__device__ __forceinline__ T tv_join(uint32_t lo, uint32_t hi)
{
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    } else {
        (void)hi;
        return __builtin_bit_cast(T, lo);
    }
}

template <typename TA, typename T, bool FORWARD, bool TRANS, int NR>
__global__ __launch_bounds__(256) void trsv_persistent_kernel(
    int n, int nrhs, const TA *__restrict__ A, int lda, T *X, int ldx, int *ticket,
    uint64_t *mail, unsigned epoch, int *abort_flag, uint64_t spin_limit)
{
    constexpr int WPV = sizeof(T) / 4;  // mailbox words per value
    __shared__ int s_b, s_abort;
    __shared__ T Xs[TV_BS][NR];
    __shared__ T Part[4][TV_BS][NR + 1];
    __shared__ T Ds[TV_BS][TV_BS + 1];  // diagonal block, row-major
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nblk = ceil_div(n, TV_BS);

    if (tid == 0) {
        s_b = atomicAdd(ticket, 1);
        s_abort = 0;
    }
    __syncthreads();
    const int t = s_b;
    if (t >= nblk) return;
    const int b = FORWARD ? t : nblk - 1 - t;
    const int rb = b * TV_BS, rows = min(TV_BS, n - rb);
    // mailbox word (block c, column j, half w, row r)
    auto mbox = [&](int c, int j, int w, int r) -> uint64_t * {
        return mail + ((((size_t)c * NR + j) * WPV + w) * TV_BS + r);
    };

    // Diagonal block staged in LDS before any wait (it does not depend on
    // other blocks); wave 0 lane r keeps 1 / U(r, r) and its right-hand side.
    for (int idx = tid; idx < TV_BS * TV_BS; idx += 256) {
        const int r = idx % TV_BS, k = idx / TV_BS;
        Ds[r][k] = (r < rows && k < rows) ? (T)A[(size_t)(rb + r) + (size_t)(rb + k) * lda] : T(0);
    }
    constexpr bool UNIT = (FORWARD != TRANS);  // L and L^T have a unit diagonal
    T rdiag = T(1);
    T rhs[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) rhs[j] = T(0);
    if (wave == 0 && lane < rows) {
        if (!UNIT) rdiag = T(1) / (T)A[(size_t)(rb + lane) * (lda + 1)];
#pragma unroll
        for (int j = 0; j < NR; j++)
            if (j < nrhs) rhs[j] = X[(size_t)(rb + lane) + (size_t)j * ldx];
    }

    T part[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) part[j] = T(0);

    // wave w covers k in [16w, 16w+16) of the 64 x 64 block A[rb.., cb..]
    // (TRANS: of the block A[cb.., rb..], read down its columns); zero
    // outside the block
    TA a[TV_PART];
    auto load_block = [&](int s) {
        const int c = FORWARD ? s : nblk - 1 - s;
        const int cb = c * TV_BS, crows = min(TV_BS, n - cb);
        const TA *Ar = TRANS ? A + (size_t)cb + (size_t)(rb + lane) * lda
                             : A + (size_t)(rb + lane) + (size_t)cb * lda;
        const size_t kstride = TRANS ? 1 : (size_t)lda;
#pragma unroll
        for (int kk = 0; kk < TV_PART; kk++) {
            const int k = wave * TV_PART + kk;
            a[kk] = (lane < rows && k < crows) ? Ar[(size_t)k * kstride] : TA(0);
        }
    };
    if (t > 0) load_block(0);

    // Off-diagonal contributions, in dependency order.
    for (int s = 0; s < t; s++) {
        const int c = FORWARD ? s : nblk - 1 - s;
        const int crows = min(TV_BS, n - c * TV_BS);
        if (wave == 0) {
            uint32_t got[NR][WPV];
            int ab = 0;
            const uint64_t t0 = wall_clock64();
            for (;;) {
                bool ok = true;
                if (lane < crows) {
#pragma unroll
                    for (int j = 0; j < NR; j++) {
                        if (j < nrhs) {
#pragma unroll
                            for (int w = 0; w < WPV; w++) {
                                const uint64_t u = __hip_atomic_load(
                                    mbox(c, j, w, lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                got[j][w] = (uint32_t)u;
                                ok = ok && (uint32_t)(u >> 32) == epoch;
                            }
                        }
                    }
                }
                if (__all(ok)) break;
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ab = 1;
                    break;
                }
                if (wall_clock64() - t0 > spin_limit) {
                    __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                // workgroups far behind the chain poll less often
                if (t - s > 4) __builtin_amdgcn_s_sleep(16);
                else __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int j = 0; j < NR; j++)
                Xs[lane][j] = (lane < crows && j < nrhs) ? tv_join<T>(got[j][0], got[j][WPV - 1])
                                                         : T(0);
            if (lane == 0) s_abort = ab;
        }
        __syncthreads();
        if (s_abort) return;  // the whole workgroup leaves together
        if (lane < rows) {
#pragma unroll
            for (int kk = 0; kk < TV_PART; kk++) {
                const T av = (T)a[kk];
                const int k = wave * TV_PART + kk;
#pragma unroll
                for (int j = 0; j < NR; j++) part[j] = fma(av, Xs[k][j], part[j]);
            }
        }
        if (s + 1 < t) load_block(s + 1);
        __syncthreads();  // Xs is reused by the next block
    }
#pragma unroll
    for (int j = 0; j < NR; j++) Part[wave][lane][j] = part[j];
    __syncthreads();

    if (wave == 0) {
        T x[NR];
#pragma unroll
        for (int j = 0; j < NR; j++) {
            const T sum = ((Part[0][lane][j] + Part[1][lane][j]) + Part[2][lane][j]) + Part[3][lane][j];
            x[j] = rhs[j] - sum;
        }
        // Column i of the triangle applied to x: element (lane, i) of the
        // diagonal block, or (i, lane) for the transposed solves.
        if (FORWARD) {
            for (int i = 0; i < rows; i++) {
                const T l = TRANS ? Ds[i][lane] : Ds[lane][i];
#pragma unroll
                for (int j = 0; j < NR; j++) {
                    if (!UNIT && lane == i) x[j] = x[j] * rdiag;
                    const T xi = lane_read(x[j], i);
                    if (lane > i) x[j] = fma(-l, xi, x[j]);
                }
            }
        } else {
            for (int i = rows - 1; i >= 0; i--) {
                const T u = TRANS ? Ds[i][lane] : Ds[lane][i];
#pragma unroll
                for (int j = 0; j < NR; j++) {
                    if (!UNIT && lane == i) x[j] = x[j] * rdiag;
                    const T xi = lane_read(x[j], i);
                    if (lane < i) x[j] = fma(-u, xi, x[j]);
                }
            }
        }
        if (lane < rows) {
            const uint64_t tag = (uint64_t)epoch << 32;
#pragma unroll
            for (int j = 0; j < NR; j++) {
                if (j < nrhs) {
#pragma unroll
                    for (int w = 0; w < WPV; w++)
                        __hip_atomic_store(mbox(b, j, w, lane), tag | tv_half(x[j], w),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    X[(size_t)(rb + lane) + (size_t)j * ldx] = x[j];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block-inverse solves with the factors of a low-precision factorization
// (the corrections of iterative refinement and GMRES-IR's preconditioner).
// The inverses of the 128 x 128 diagonal blocks of L and U are formed once
// per factorization, so a step of the dependency chain is two products with
// blocks already in registers instead of 128 dependent substitution steps:
//   x_b = D_b^{-1} (rhs_b - sum_{c before b} T[b, c] x_c).
constexpr int TB_BS = 128, TB_KH = TB_BS / 2, TB_INV_COLS = 64;

// inv holds plane 0 = the inverses of L's unit-lower diagonal blocks and
// plane 1 = those of U's upper diagonal blocks, FP64, block b at
// inv + (plane * nblk + b) * 128 * 128, column-major (element (r, k) at
// k * 128 + r), zero outside the triangle; rows past n are the identity's.
// Workgroup (b, half, plane): one wave, lane j forms column 64 half + j by
// substitution against e_j in FP64, the block staged in LDS (all lanes read
// the same element: a broadcast) and the column in LDS (each lane its own).
template <typename TA>
__global__ __launch_bounds__(TB_INV_COLS) void trsv_diag_inverse_kernel(int n, const TA *__restrict__ A,
                                                                       int lda, double *__restrict__ inv)
{
    __shared__ TA Ds[TB_BS][TB_BS + 1];
    __shared__ double Xs[TB_BS][TB_INV_COLS];
    const int b = blockIdx.x, half = blockIdx.y, plane = blockIdx.z;
    const int nblk = gridDim.x, rb = b * TB_BS, rows = min(TB_BS, n - rb);
    const int lane = threadIdx.x, j = half * TB_INV_COLS + lane;
    for (int idx = lane; idx < TB_BS * TB_BS; idx += TB_INV_COLS) {
        const int r = idx % TB_BS, k = idx / TB_BS;
        TA v = (r == k) ? TA(1) : TA(0);
        if (r < rows && k < rows) v = A[(size_t)(rb + r) + (size_t)(rb + k) * lda];
        Ds[r][k] = v;
    }
    __syncthreads();
    if (plane == 0) {
        // unit lower: x_i = e_j(i) - sum_{k < i} L(i, k) x_k
        for (int i = 0; i < TB_BS; i++) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) s = fma(-(double)Ds[i][k], Xs[k][lane], s);
            Xs[i][lane] = s;
        }
    } else {
        // upper: x_i = (e_j(i) - sum_{k > i} U(i, k) x_k) / U(i, i)
        for (int i = TB_BS - 1; i >= 0; i--) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = i + 1; k < TB_BS; k++) s = fma(-(double)Ds[i][k], Xs[k][lane], s);
            Xs[i][lane] = s / (double)Ds[i][i];
        }
    }
    double *out = inv + ((size_t)plane * nblk + b) * TB_BS * TB_BS + (size_t)j * TB_BS;
    for (int i = 0; i < TB_BS; i++) out[i] = Xs[i][lane];
}

// X <- L^{-1} X (FORWARD) or U^{-1} X for the FP32 factors at A with the
// block inverses at inv (plane 0 / 1), nrhs <= NR <= 16, T the arithmetic.
// Same ticket / epoch-tagged mailbox protocol as trsv_persistent_kernel, with
// 128-row blocks: thread (r, h) = (tid % 128, tid / 128) owns row r and the
// k-half h of every block it multiplies; waves 0 and 1 poll the 128 rows of
// the block they wait for.
template <typename T, bool FORWARD, int NR>
__global__ __launch_bounds__(256) void trsv_blockinv_kernel(
    int n, int nrhs, const float *__restrict__ A, int lda, const double *__restrict__ inv, T *X,
    int ldx, int *ticket, uint64_t *mail, unsigned epoch, int *abort_flag, uint64_t spin_limit)
{
    constexpr int WPV = sizeof(T) / 4;
    __shared__ int s_b, s_abort;
    __shared__ T Xs[TB_BS][NR];
    __shared__ T Ph[2][TB_BS][NR + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = tid % TB_BS, h = tid / TB_BS;
    const int nblk = ceil_div(n, TB_BS);
    if (tid == 0) {
        s_b = atomicAdd(ticket, 1);
        s_abort = 0;
    }
    __syncthreads();
    const int t = s_b;
    if (t >= nblk) return;
    const int b = FORWARD ? t : nblk - 1 - t;
    const int rb = b * TB_BS, rows = min(TB_BS, n - rb);
    auto mbox = [&](int c, int j, int w, int row) -> uint64_t * {
        return mail + ((((size_t)c * NR + j) * WPV + w) * TB_BS + row);
    };

    // the diagonal block's inverse, row r, columns of half h (before any wait)
    T dv[TB_KH];
    {
        const double *Di = inv + ((size_t)(FORWARD ? 0 : 1) * nblk + b) * TB_BS * TB_BS;
#pragma unroll
        for (int kk = 0; kk < TB_KH; kk++) dv[kk] = (T)Di[(size_t)(h * TB_KH + kk) * TB_BS + r];
    }
    T rhs[NR], part[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) {
        rhs[j] = (h == 0 && r < rows && j < nrhs) ? X[(size_t)(rb + r) + (size_t)j * ldx] : T(0);
        part[j] = T(0);
    }
    float a[TB_KH];
    auto load_block = [&](int s) {
        const int c = FORWARD ? s : nblk - 1 - s;
        const int cb = c * TB_BS, crows = min(TB_BS, n - cb);
        const float *Ar = A + (size_t)(rb + r) + (size_t)(cb + h * TB_KH) * lda;
#pragma unroll
        for (int kk = 0; kk < TB_KH; kk++)
            a[kk] = (r < rows && h * TB_KH + kk < crows) ? Ar[(size_t)kk * lda] : 0.f;
    };
    if (t > 0) load_block(0);

    for (int s = 0; s < t; s++) {
        const int c = FORWARD ? s : nblk - 1 - s;
        const int crows = min(TB_BS, n - c * TB_BS);
        if (wave < 2) {
            const int row = wave * 64 + lane;
            uint32_t got[NR][WPV];
            int ab = 0;
            const uint64_t t0 = wall_clock64();
            for (;;) {
                bool ok = true;
                if (row < crows) {
#pragma unroll
                    for (int j = 0; j < NR; j++) {
                        if (j < nrhs) {
#pragma unroll
                            for (int w = 0; w < WPV; w++) {
                                const uint64_t u = __hip_atomic_load(
                                    mbox(c, j, w, row), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                got[j][w] = (uint32_t)u;
                                ok = ok && (uint32_t)(u >> 32) == epoch;
                            }
                        }
                    }
                }
                if (__all(ok)) break;
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    ab = 1;
                    break;
                }
                if (wall_clock64() - t0 > spin_limit) {
                    __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                if (t - s > 4) __builtin_amdgcn_s_sleep(16);
                else __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int j = 0; j < NR; j++)
                Xs[row][j] = (row < crows && j < nrhs) ? tv_join<T>(got[j][0], got[j][WPV - 1]) : T(0);
            if (ab) s_abort = 1;
        }
        __syncthreads();
        if (s_abort) return;  // the whole workgroup leaves together
#pragma unroll
        for (int kk = 0; kk < TB_KH; kk++) {
            const T av = (T)a[kk];
#pragma unroll
            for (int j = 0; j < NR; j++) part[j] = fma(av, Xs[h * TB_KH + kk][j], part[j]);
        }
        if (s + 1 < t) load_block(s + 1);
        __syncthreads();  // Xs is reused by the next block
    }
    // y = rhs - (part_0 + part_1), then x = D^{-1} y as two half products
#pragma unroll
    for (int j = 0; j < NR; j++) Ph[h][r][j] = part[j];
    __syncthreads();
    if (h == 0) {
#pragma unroll
        for (int j = 0; j < NR; j++) Xs[r][j] = r < rows ? rhs[j] - (Ph[0][r][j] + Ph[1][r][j]) : T(0);
    }
    __syncthreads();
    T q[NR];
#pragma unroll
    for (int j = 0; j < NR; j++) q[j] = T(0);
#pragma unroll
    for (int kk = 0; kk < TB_KH; kk++) {
        const T d = dv[kk];
#pragma unroll
        for (int j = 0; j < NR; j++) q[j] = fma(d, Xs[h * TB_KH + kk][j], q[j]);
    }
#pragma unroll
    for (int j = 0; j < NR; j++) Ph[h][r][j] = q[j];
    __syncthreads();
    if (h == 0 && r < rows) {
        const uint64_t tag = (uint64_t)epoch << 32;
#pragma unroll
        for (int j = 0; j < NR; j++) {
            if (j < nrhs) {
                const T x = Ph[0][r][j] + Ph[1][r][j];
#pragma unroll
                for (int w = 0; w < WPV; w++)
                    __hip_atomic_store(mbox(b, j, w, r), tag | tv_half(x, w), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                X[(size_t)(rb + r) + (size_t)j * ldx] = x;
            }
        }
    }
}

}  // namespace

void trsv_block_inverses(int n, const float *A, int lda, double *inv, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(trsv_diag_inverse_kernel<float>, dim3(ceil_div(n, TB_BS), 2, 2),
                       dim3(TB_INV_COLS), 0, s, n, A, lda, inv);
    SLA_LAUNCH_CHECK();
}

template <typename T>
void trsv_blockinv(bool lower, int n, int nrhs, const float *A, int lda, const double *inv, T *X,
                   int ldx, TrsvWork &tw, hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return;
    if (nrhs > TV_NR) throw std::logic_error("trsv_blockinv: at most 16 right-hand sides per call");
    const int nblk = ceil_div(n, TB_BS);
    if ((size_t)nblk * TB_BS * TV_NR * 2 > tw.mail_words)
        throw std::logic_error("trsv_blockinv: mailbox too small for this order");
    fill_int(tw.ticket, 1, 0, s);
    const int epoch = ++tw.epoch;
#define SLA_TBI(FW, NRV)                                                                          \
    hipLaunchKernelGGL((trsv_blockinv_kernel<T, FW, NRV>), dim3(nblk), dim3(256), 0, s, n, nrhs, \
                       A, lda, inv, X, ldx, tw.ticket, tw.mail, (unsigned)epoch, tw.abort_flag,  \
                       tw.spin_limit)
#define SLA_TBI_NR(FW)                          \
    do {                                        \
        if (nrhs <= 1) SLA_TBI(FW, 1);          \
        else if (nrhs <= 2) SLA_TBI(FW, 2);     \
        else if (nrhs <= 4) SLA_TBI(FW, 4);     \
        else if (nrhs <= 8) SLA_TBI(FW, 8);     \
        else SLA_TBI(FW, 16);                   \
    } while (0)
    if (lower) SLA_TBI_NR(true);
    else SLA_TBI_NR(false);
#undef SLA_TBI_NR
#undef SLA_TBI
    SLA_LAUNCH_CHECK();
}
template void trsv_blockinv<float>(bool, int, int, const float *, int, const double *, float *,
                                   int, TrsvWork &, hipStream_t);
template void trsv_blockinv<double>(bool, int, int, const float *, int, const double *, double *,
                                    int, TrsvWork &, hipStream_t);

template <typename T>
void trsm_llnu(int m, int n, const T *L, int ldl, T *B, int ldb, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    if (m > TS_MAXM) throw std::logic_error("trsm_llnu: at most 256 rows per call");
    hipLaunchKernelGGL(trsm_llnu_strip_kernel<T>, dim3(ceil_div(n, TS_CW)), dim3(256), 0, s, m, n,
                       L, ldl, B, ldb);
    SLA_LAUNCH_CHECK();
}
template void trsm_llnu<double>(int, int, const double *, int, double *, int, hipStream_t);
template void trsm_llnu<float>(int, int, const float *, int, float *, int, hipStream_t);

// ---------------------------------------------------------------------------
// trsm_llnu_blocked: the same solve for m <= 512 rows in one launch (plus one
// for the inverses, which a second solve with the same L11 reuses), the
// products on MFMA.  L11 is taken in 32-row blocks; the inverse of each unit
// lower 32 x 32 diagonal block is formed by trsm_diag_inv_kernel (thread =
// element, forward substitution against the identity, one barrier per row;
// a block cut by m is completed with the identity).  Each workgroup (8
// waves) owns 16 columns of B, held in LDS, and for block k
//   X_k = inv(L_kk) B_k                     (waves 0, 1: a 16-row tile each)
//   B_i -= L_ik X_k  for the rows below     (16-row tiles, two per wave at a
//                                            time so 16 loads are in flight)
// with v_mfma_f64_16x16x4f64 (FP32: v_mfma_f32_16x16x4f32), L read from
// global memory (L2-resident across the workgroups), X_k from LDS; the next
// block's inverse operands are loaded during the current block's update.
namespace {

constexpr int TB_M = 512, TB_NC = 16, TB_LD = TB_NC + 1, TB_T = 512;
typedef double bd4_t __attribute__((ext_vector_type(4)));
typedef float bf4_t __attribute__((ext_vector_type(4)));

template <typename T>
__global__ __launch_bounds__(1024) void trsm_diag_inv_kernel(int m, const T *__restrict__ L,
                                                             int ldl, T *__restrict__ inv)
{
    // thread (i, j) holds element (i, j) of the inverse; step k subtracts
    // D_ik x_kj from the rows below k (forward substitution, column-parallel)
    __shared__ T Ds[32][33], Xs[32][33];
    const int b = blockIdx.x, tid = threadIdx.x, i = tid % 32, j = tid / 32, r0 = 32 * b;
    Ds[i][j] = (r0 + i < m && r0 + j < m && i > j) ? L[(size_t)(r0 + i) + (size_t)(r0 + j) * ldl]
                                                  : T(0);
    T x = (i == j) ? T(1) : T(0);
    Xs[i][j] = x;
    for (int k = 0; k < 31; k++) {
        __syncthreads();
        if (i > k) {
            x = fma(-Ds[i][k], Xs[k][j], x);
            if (i == k + 1) Xs[i][j] = x;  // row k + 1 is final
        }
    }
    inv[(size_t)b * 1024 + (size_t)j * 32 + i] = x;
}

template <typename T>
__device__ __forceinline__ auto mfma16x4(T a, T b, typename std::conditional<sizeof(T) == 8, bd4_t,
                                                                            bf4_t>::type c)
{
    if constexpr (sizeof(T) == 8) return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of accumulator element r of lane (lr, lk) within a 16 x 16 tile
template <typename T>
__device__ __forceinline__ int mfma_row(int lk, int r)
{
    return sizeof(T) == 8 ? lk + 4 * r : 4 * lk + r;
}

template <typename T>
__global__ __launch_bounds__(TB_T) void trsm_llnu_blocked_kernel(int m, int n,
                                                                 const T *__restrict__ L, int ldl,
                                                                 const T *__restrict__ inv,
                                                                 T *__restrict__ B, int ldb)
{
    using acc_t = typename std::conditional<sizeof(T) == 8, bd4_t, bf4_t>::type;
    constexpr int NW = TB_T / 64;
    __shared__ T Xs[TB_M][TB_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int c0 = blockIdx.x * TB_NC;
    const int nblk = (m + 31) / 32, m32 = 32 * nblk;
    // B strip into LDS (rows >= m and columns >= n as zeros)
    for (int idx = tid; idx < m32 * TB_NC; idx += TB_T) {
        const int r = idx % m32, c = idx / m32;
        Xs[r][c] = (r < m && c0 + c < n) ? B[(size_t)r + (size_t)(c0 + c) * ldb] : T(0);
    }
    // inverse operands of block 0 (waves 0, 1)
    T iva[8];
    if (wave < 2) {
#pragma unroll
        for (int s = 0; s < 8; s++) iva[s] = inv[(size_t)(4 * s + lk) * 32 + 16 * wave + lr];
    }
    __syncthreads();
    for (int k = 0; k < nblk; k++) {
        const int kr = 32 * k;
        // X_k = inv(L_kk) B_k: waves 0, 1 take rows kr + 16 wave ..
        acc_t d = acc_t{0, 0, 0, 0};
        if (wave < 2) {
#pragma unroll
            for (int s = 0; s < 8; s++) d = mfma16x4<T>(iva[s], Xs[kr + 4 * s + lk][lr], d);
            if (k + 1 < nblk) {  // next block's inverse, in flight during this block's update
                const T *iv = inv + (size_t)(k + 1) * 1024;
#pragma unroll
                for (int s = 0; s < 8; s++) iva[s] = iv[(size_t)(4 * s + lk) * 32 + 16 * wave + lr];
            }
        }
        __syncthreads();  // every read of B_k is done
        if (wave < 2) {
#pragma unroll
            for (int r = 0; r < 4; r++) Xs[kr + 16 * wave + mfma_row<T>(lk, r)][lr] = d[r];
        }
        __syncthreads();  // X_k is final
        const int ntile = (m32 - kr - 32) / 16;  // 16-row tiles below block k
        if (ntile <= 0) break;
        // -X_k as the B operand of every tile of this block
        T xb[8];
#pragma unroll
        for (int s = 0; s < 8; s++) xb[s] = -Xs[kr + 4 * s + lk][lr];
        // tiles t and t + NW together: 16 loads of L in flight per lane
        for (int t = wave; t < ntile; t += 2 * NW) {
            T a[2][8];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int row = kr + 32 + 16 * (t + h * NW) + lr;
                const T *lp = L + (size_t)row + (size_t)kr * ldl;
                const bool in = t + h * NW < ntile && row < m;
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int col = 4 * s + lk;
                    a[h][s] = (in && kr + col < m) ? lp[(size_t)col * ldl] : T(0);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (t + h * NW >= ntile) break;
                const int r0 = kr + 32 + 16 * (t + h * NW);
                acc_t c;
#pragma unroll
                for (int r = 0; r < 4; r++) c[r] = Xs[r0 + mfma_row<T>(lk, r)][lr];
#pragma unroll
                for (int s = 0; s < 8; s++) c = mfma16x4<T>(a[h][s], xb[s], c);
#pragma unroll
                for (int r = 0; r < 4; r++) Xs[r0 + mfma_row<T>(lk, r)][lr] = c[r];
            }
        }
        __syncthreads();  // block k + 1 is updated
    }
    for (int idx = tid; idx < m * TB_NC; idx += TB_T) {
        const int r = idx % m, c = idx / m;
        if (c0 + c < n) B[(size_t)r + (size_t)(c0 + c) * ldb] = Xs[r][c];
    }
}

}  // namespace

template <typename T>
void trsm_llnu_blocked(int m, int n, const T *L, int ldl, T *B, int ldb, T *inv, bool inv_ready,
                       hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    if (m > TB_M) throw std::logic_error("trsm_llnu_blocked: at most 512 rows per call");
    const int nblk = (m + 31) / 32;
    if (!inv_ready) {
        hipLaunchKernelGGL(trsm_diag_inv_kernel<T>, dim3(nblk), dim3(1024), 0, s, m, L, ldl, inv);
        SLA_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(trsm_llnu_blocked_kernel<T>, dim3(ceil_div(n, TB_NC)), dim3(TB_T), 0, s, m,
                       n, L, ldl, (const T *)inv, B, ldb);
    SLA_LAUNCH_CHECK();
}
template void trsm_llnu_blocked<double>(int, int, const double *, int, double *, int, double *,
                                        bool, hipStream_t);
template void trsm_llnu_blocked<float>(int, int, const float *, int, float *, int, float *, bool,
                                       hipStream_t);

template <typename TA, typename T>
void trsv_persistent(TrsvOp op, int n, int nrhs, const TA *A, int lda, T *X, int ldx,
                     TrsvWork &tw, hipStream_t s)
{
    if (n <= 0 || nrhs <= 0) return;
    if (nrhs > TV_NR) throw std::logic_error("trsv_persistent: at most 16 right-hand sides per call");
    const int nblk = ceil_div(n, TV_BS);
    if ((size_t)nblk * TV_BS * TV_NR * 2 > tw.mail_words)
        throw std::logic_error("trsv_persistent: mailbox too small for this order");
    fill_int(tw.ticket, 1, 0, s);
    const int epoch = ++tw.epoch;
#define SLA_TRSV(FW, TR, NRV)                                                                  \
    hipLaunchKernelGGL((trsv_persistent_kernel<TA, T, FW, TR, NRV>), dim3(nblk), dim3(256), 0, \
                       s, n, nrhs, A, lda, X, ldx, tw.ticket, tw.mail, (unsigned)epoch,        \
                       tw.abort_flag, tw.spin_limit)
#define SLA_TRSV_NR(FW, TR)                           \
    do {                                              \
        if (nrhs <= 1) SLA_TRSV(FW, TR, 1);           \
        else if (nrhs <= 2) SLA_TRSV(FW, TR, 2);      \
        else if (nrhs <= 4) SLA_TRSV(FW, TR, 4);      \
        else if (nrhs <= 8) SLA_TRSV(FW, TR, 8);      \
        else SLA_TRSV(FW, TR, 16);                    \
    } while (0)
    switch (op) {
    case TrsvOp::L: SLA_TRSV_NR(true, false); break;
    case TrsvOp::U: SLA_TRSV_NR(false, false); break;
    case TrsvOp::UT: SLA_TRSV_NR(true, true); break;
    case TrsvOp::LT: SLA_TRSV_NR(false, true); break;
    }
#undef SLA_TRSV_NR
#undef SLA_TRSV
    SLA_LAUNCH_CHECK();
}
template void trsv_persistent<double, double>(TrsvOp, int, int, const double *, int, double *, int,
                                              TrsvWork &, hipStream_t);
template void trsv_persistent<float, float>(TrsvOp, int, int, const float *, int, float *, int,
                                            TrsvWork &, hipStream_t);
template void trsv_persistent<float, double>(TrsvOp, int, int, const float *, int, double *, int,
                                             TrsvWork &, hipStream_t);

}  // namespace sla
