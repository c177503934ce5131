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

}  // namespace

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
