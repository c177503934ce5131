// batched.hip -- many independent small systems A_s X_s = B_s (n_s <= 1024,
// nrhs_s <= 16, uneven sizes) factored and solved in ONE persistent launch.
//
// Each system follows the LU path of a 1 x 1 grid: partial pivoting with the
// first maximum |a| of each column (IDAMAX), rows interchanged across the
// whole matrix, the pivot column scaled by the reciprocal of the pivot
// (SRC/pdgetf2.f:206-237), a zero pivot recorded as INFO = column and the
// factorization continued (pdgetrf.f:297-302); then PDGETRS's interchanges
// and two triangular solves (pdgetrs.f:255-266), skipped when INFO > 0
// (pdgesv.f:231).
//
// Work distribution: workgroups (256 threads) take systems from a device
// queue, largest first (the host sorts the sizes once), until the queue is
// empty; every workgroup leaves when the ticket passes the batch size.
// Inside a system the factorization is blocked right-looking with W-column
// panels:
//   panel    columns kb .. kb+W-1: while the panel has at most 256 RPT rows
//            (RPT = 2 for FP64, 4 for FP32) it is held in registers and a
//            column costs one workgroup barrier (panel_regs); taller panels
//            run unblocked in memory (pivot search as a wave-shuffle argmax
//            folded across the four waves in LDS, row interchange over all n
//            columns, reciprocal scaling, rank-1 update, two barriers per
//            column);
//   U12      L11^{-1} A12, one thread per column of A12;
//   update   A22 -= L21 U12 in 64 x 64 tiles, L21 / U12 staged through LDS,
//            the products on the 16 x 16 x 4 MFMA of the storage precision.
// The solves take 64-row blocks: one wave solves the diagonal block in
// registers (lane = row, values broadcast by lane reads), then all four
// waves update the rows below (above) with the 64 new values.
#include "sla_common.h"
#include "sla_handle.h"
#include "sla_batched.h"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace sla {

namespace {

constexpr int BT = 256;       // threads per workgroup
constexpr int UT = 64;        // update tile edge

// Kernel variants for comparisons: SLA_BATCHED_FLAGS (environment) sets the
// flags of factor_system (default 3: register panels and MFMA updates).
int batched_flags()
{
    static const int f = [] {
        const char *e = getenv("SLA_BATCHED_FLAGS");
        return e ? (atoi(e) & 3) : 3;
    }();
    return f;
}

struct SysDesc {
    int n, nrhs, lda, ldb;
};

template <typename T>
__device__ __forceinline__ void argmax_fold(T &v, int &i, T v2, int i2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}
// The same carrying the signed value at the winning row.
template <typename T>
__device__ __forceinline__ void argmax_fold3(T &v, int &i, T &sv, T v2, int i2, T sv2)
{
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
        sv = sv2;
    }
}

// LDS of the register-resident panel: per column parity, the candidate row
// of every wave (slots 0..3) and the row at the diagonal position (slot 4),
// the wave maxima and their positions; the panel's pivots.
template <typename T, int W>
struct RegPanelLds {
    T rows[2][5][W];
    T val[2][4];
    int pos[2][4];
    int piv[W];
};

// Panel columns kb .. kb+w-1 of rows kb .. n-1 held in registers while they
// are factored: thread t holds positions kb + t + 256 q (q < RPT), so this
// path serves panels of at most 256 RPT rows.  Per column j: every thread
// finds its largest |a| among the positions >= j (lowest position on ties),
// a wave argmax, the owner of each wave's winner posts its whole panel row
// and the owner of position j its row, ONE workgroup barrier, then every
// thread folds the four posts in the same order: the owner of the pivot's
// position takes the row from position j and the owner of position j the
// pivot row (the interchange within the panel), the rows below j are scaled
// by the reciprocal of the pivot and updated with the pivot row.  The
// posts are double-buffered by column parity, so no second barrier is
// needed.  The operations on every element are those of the unblocked
// panel below (same interchange, reciprocal scaling, FMA), so both give the
// same bits.  The interchanges reach the other columns after the panel,
// replayed in order column by column.
template <typename T, int W, int RPT>
__device__ void panel_regs(int n, int kb, int w, T *__restrict__ a, int lda, int *__restrict__ piv,
                           int &info, RegPanelLds<T, W> &sl)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    T v[RPT][W];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = kb + tid + BT * q;
#pragma unroll
        for (int c = 0; c < W; c++)
            v[q][c] = (r < n && c < w) ? a[(size_t)r + (size_t)(kb + c) * lda] : T(0);
    }
#pragma unroll
    for (int jj = 0; jj < W; jj++) {
        if (jj < w) {  // uniform
        const int j = kb + jj, pb = jj & 1;
        // the largest |a| of column j among this thread's positions >= j
        T best = T(-1);
        int bi = INT_MAX;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = kb + tid + BT * q;
            const T x = fabs(v[q][jj]);
            if (r >= j && r < n && x > best) {  // a NaN never wins
                best = x;
                bi = r;
            }
        }
        for (int o = 32; o > 0; o >>= 1) argmax_fold(best, bi, __shfl_xor(best, o), __shfl_xor(bi, o));
        if (bi != INT_MAX && (bi - kb) % BT == tid) {  // this thread holds the wave's winner
            const int qw = (bi - kb) / BT;
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == qw)
#pragma unroll
                    for (int c = 0; c < W; c++) sl.rows[pb][wave][c] = v[q][c];
        }
        if (tid == jj)  // position j is (thread jj, q = 0)
#pragma unroll
            for (int c = 0; c < W; c++) sl.rows[pb][4][c] = v[0][c];
        if (lane == 0) {
            sl.val[pb][wave] = best;
            sl.pos[pb][wave] = bi;
        }
        __syncthreads();
        T fv = sl.val[pb][0];
        int fi = sl.pos[pb][0], fw = 0;
#pragma unroll
        for (int q = 1; q < 4; q++) {
            const T v2 = sl.val[pb][q];
            const int i2 = sl.pos[pb][q];
            if (v2 > fv || (v2 == fv && i2 < fi)) {
                fv = v2;
                fi = i2;
                fw = q;
            }
        }
        // no comparable entry (all NaN): keep row j
        const int p = fi < n ? fi : j;
        const T *prow = fi < n ? sl.rows[pb][fw] : sl.rows[pb][4];
        const T pv = prow[jj];
        if (tid == 0) {
            piv[j] = p + 1;
            sl.piv[jj] = p;
        }
        if (pv != T(0)) {
            if (p != j) {
                if ((p - kb) % BT == tid) {
                    const int qp = (p - kb) / BT;
#pragma unroll
                    for (int q = 0; q < RPT; q++)
                        if (q == qp)
#pragma unroll
                            for (int c = 0; c < W; c++) v[q][c] = sl.rows[pb][4][c];
                }
                if (tid == jj)
#pragma unroll
                    for (int c = 0; c < W; c++) v[0][c] = prow[c];
            }
            const T rcp = T(1) / pv;
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                const int r = kb + tid + BT * q;
                if (r > j && r < n) v[q][jj] = v[q][jj] * rcp;
            }
        } else if (info == 0) {
            info = j + 1;
        }
        // rank-1 update of the panel's columns right of j
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = kb + tid + BT * q;
            if (r > j && r < n) {
                const T l = v[q][jj];
                // constant bounds, so the inner loop unrolls before the outer
#pragma unroll
                for (int c = 1; c < W; c++)
                    if (c > jj && c < w) v[q][c] = fma(-l, prow[c], v[q][c]);
            }
        }
        }  // jj < w
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = kb + tid + BT * q;
        if (r < n)
#pragma unroll
            for (int c = 0; c < W; c++)
                if (c < w) a[(size_t)r + (size_t)(kb + c) * lda] = v[q][c];
    }
    __syncthreads();  // sl.piv complete
    // the panel's interchanges, in order, on every other column
    for (int c = tid; c < n; c += BT) {
        if (c >= kb && c < kb + w) continue;
        T *cc = a + (size_t)c * lda;
        for (int jj = 0; jj < w; jj++) {
            const int p = sl.piv[jj];
            if (p != kb + jj) {
                const T t = cc[kb + jj];
                cc[kb + jj] = cc[p];
                cc[p] = t;
            }
        }
    }
}

// C[rows x cols] -= Ls[0:rows, 0:w] Us[0:w, 0:cols] for one 64 x 64 tile on
// the 16 x 16 x 4 MFMA of the storage precision: wave v owns rows
// 32 (v & 1) .., columns 32 (v >> 1) .. as 2 x 2 blocks.  w is a multiple of
// 4 (every update follows a full panel of width W in {8, 16, 32}).
// D layouts: f64 16x16x4, element r of lane l = D[l/16 + 4 r][l % 16];
// f32 16x16x4, element r of lane l = D[4 (l/16) + r][l % 16].
typedef double bd4_t __attribute__((ext_vector_type(4)));
typedef float bf4_t __attribute__((ext_vector_type(4)));

template <typename T, int W>
__device__ __forceinline__ void tile_update_mfma(int w, int rows, int cols, const T (*Ls)[W + 1],
                                                 const T (*Us)[UT + 1], T *__restrict__ C, int ldc)
{
    using acc_t = typename std::conditional<sizeof(T) == 8, bd4_t, bf4_t>::type;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32, lr = lane & 15, lk = lane >> 4;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = acc_t{0, 0, 0, 0};
    for (int k0 = 0; k0 < w; k0 += 4) {
        T af[2], bf[2];
#pragma unroll
        for (int i = 0; i < 2; i++) af[i] = Ls[wr + 16 * i + lr][k0 + lk];
#pragma unroll
        for (int j = 0; j < 2; j++) bf[j] = Us[k0 + lk][wc + 16 * j + lr];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if constexpr (sizeof(T) == 8)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
                else
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int c = wc + 16 * j + lr;
        if (c >= cols) continue;
        T *cc = C + (size_t)c * ldc;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wr + 16 * i + (sizeof(T) == 8 ? lk + 4 * r : 4 * lk + r);
                if (row < rows) cc[row] -= acc[i][j][r];
            }
    }
}

// The same with 4 x 4 outputs per thread on vector FMAs (thread t: rows
// 4 (t % 16) .., columns 4 (t / 16) ..).
template <typename T, int W>
__device__ __forceinline__ void tile_update_valu(int w, int rows, int cols, const T (*Ls)[W + 1],
                                                 const T (*Us)[UT + 1], T *__restrict__ C, int ldc)
{
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    T acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = T(0);
    for (int k = 0; k < w; k++) {
        T l[4], u[4];
#pragma unroll
        for (int x = 0; x < 4; x++) l[x] = Ls[4 * tr + x][k];
#pragma unroll
        for (int y = 0; y < 4; y++) u[y] = Us[k][4 * tc + y];
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) acc[x][y] = fma(l[x], u[y], acc[x][y]);
    }
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const int c = 4 * tc + y;
        if (c < cols) {
            T *cc = C + (size_t)c * ldc;
#pragma unroll
            for (int x = 0; x < 4; x++)
                if (4 * tr + x < rows) cc[4 * tr + x] -= acc[x][y];
        }
    }
}

// flags: bit 0 -- panels of at most 256 RPT rows in registers (panel_regs);
// bit 1 -- the trailing-update tiles on MFMA (else vector FMAs).
template <typename T, int W>
__device__ void factor_system(int n, T *__restrict__ a, int lda, int *__restrict__ piv, int &info,
                              T (*Ls)[W + 1], T (*Us)[UT + 1], T (*L11s)[W + 1], T *s_v, int *s_i,
                              RegPanelLds<T, W> &rp, int flags)
{
    constexpr int RPT = sizeof(T) == 8 ? 2 : 4;  // rows per thread of the register panel
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int kb = 0; kb < n; kb += W) {
        const int w = min(W, n - kb);
        if ((flags & 1) && n - kb <= BT * RPT) {
            panel_regs<T, W, RPT>(n, kb, w, a, lda, piv, info, rp);
        } else {
        // ---- panel (unblocked), columns kb .. kb+w-1.  Thread t owns rows
        // j+1+t, j+1+t+256, ... of every step, so the rank-1 update of
        // column j and the pivot search of column j+1 run in one pass over
        // those rows with no barrier between them.  The search carries the
        // signed value of the winning row, and the owner of row j (thread 0)
        // publishes A(j, j) before the barrier, so after it no thread reads
        // the two elements that the interchange then rewrites; the
        // interchange of the other columns and the scaling of column j touch
        // disjoint data: two workgroup barriers per column.
        T best = T(-1), bsv = T(0);
        int bi = INT_MAX;
        {
            const T *col = a + (size_t)kb * lda;
            for (int i = kb + tid; i < n; i += BT) {
                const T x = col[i], v = fabs(x);
                if (v > best) {
                    best = v;
                    bi = i;
                    bsv = x;
                }
            }
            if (tid == 0) s_v[4] = col[kb];
        }
        for (int j = kb; j < kb + w; j++) {
            T *cj = a + (size_t)j * lda;
            for (int o = 32; o > 0; o >>= 1)
                argmax_fold3(best, bi, bsv, __shfl_xor(best, o), __shfl_xor(bi, o), __shfl_xor(bsv, o));
            if (lane == 0) {
                s_v[wave] = best;
                s_i[wave] = bi;
                s_v[5 + wave] = bsv;
            }
            __syncthreads();
            best = s_v[0];
            bi = s_i[0];
            bsv = s_v[5];
            for (int q = 1; q < 4; q++) argmax_fold3(best, bi, bsv, s_v[q], s_i[q], s_v[5 + q]);
            const T ajj = s_v[4];
            const int p = bi < n ? bi : j;  // no comparable entry (all NaN): keep row j
            const T pv = bi < n ? bsv : ajj;
            if (tid == 0) piv[j] = p + 1;
            if (pv != T(0)) {
                if (p != j)
                    for (int c = tid; c < n; c += BT) {
                        if (c == j) continue;
                        T *cc = a + (size_t)c * lda;
                        const T t = cc[j];
                        cc[j] = cc[p];
                        cc[p] = t;
                    }
                const T r = T(1) / pv;
                // column j: rows j+1.. scaled (row p receives the old A(j, j))
                for (int i = j + 1 + tid; i < n; i += BT) cj[i] = (i == p ? ajj : cj[i]) * r;
                if (tid == 0) cj[j] = pv;
            } else if (info == 0) {
                info = j + 1;
            }
            __syncthreads();  // row j (all columns) and column j are final
            // rank-1 update of the panel's columns right of j, fused with the
            // pivot search of column j + 1 over the same rows
            best = T(-1);
            bi = INT_MAX;
            bsv = T(0);
            const int cend = kb + w;
            for (int i = j + 1 + tid; i < n; i += BT) {
                const T l = cj[i];
                for (int c = j + 1; c < cend; c++) {
                    T *cc = a + (size_t)c * lda;
                    cc[i] = fma(-l, cc[j], cc[i]);
                }
                if (j + 1 < cend) {
                    const T x = a[(size_t)(j + 1) * lda + i], v = fabs(x);
                    if (v > best) {
                        best = v;
                        bi = i;
                        bsv = x;
                    }
                }
            }
            // row j+1 is owned by thread 0 in this pass: A(j+1, j+1) is final
            if (tid == 0 && j + 1 < cend && j + 1 < n) s_v[4] = a[(size_t)(j + 1) * (lda + 1)];
        }
        }  // panel
        __syncthreads();  // the panel is final before U12 reads L11
        const int c0 = kb + w;
        if (c0 >= n) break;
        // ---- U12 = L11^{-1} A12 and A22 -= L21 U12, by 64-column chunks of
        // the trailing columns: the chunk of A12 is staged in LDS once,
        // solved there against L11 (one thread per column, the FMA sequence
        // of a substitution) and written back as U12, then every 64-row tile
        // of A22 below it is updated with it (L21 tiles staged in LDS, the
        // products on MFMA or vector FMAs)
        const int m2 = n - c0;
        for (int idx = tid; idx < w * w; idx += BT) {
            const int r = idx % w, k = idx / w;
            L11s[r][k] = a[(size_t)(kb + r) + (size_t)(kb + k) * lda];
        }
        for (int u0 = 0; u0 < m2; u0 += UT) {
            const int cols = min(UT, m2 - u0);
            for (int idx = tid; idx < UT * w; idx += BT) {
                const int k = idx % w, c = idx / w;
                Us[k][c] = c < cols ? a[(size_t)(kb + k) + (size_t)(c0 + u0 + c) * lda] : T(0);
            }
            __syncthreads();
            if (tid < cols) {
                for (int r = 0; r < w; r++) {
                    const T x = Us[r][tid];
                    for (int r2 = r + 1; r2 < w; r2++) Us[r2][tid] = fma(-L11s[r2][r], x, Us[r2][tid]);
                }
            }
            __syncthreads();
            for (int idx = tid; idx < UT * w; idx += BT) {
                const int k = idx % w, c = idx / w;
                if (c < cols) a[(size_t)(kb + k) + (size_t)(c0 + u0 + c) * lda] = Us[k][c];
            }
            for (int t0 = 0; t0 < m2; t0 += UT) {
                const int rows = min(UT, m2 - t0);
                for (int idx = tid; idx < UT * w; idx += BT) {
                    const int i = idx % UT, k = idx / UT;
                    Ls[i][k] = i < rows ? a[(size_t)(c0 + t0 + i) + (size_t)(kb + k) * lda] : T(0);
                }
                __syncthreads();
                if (flags & 2)
                    tile_update_mfma<T, W>(w, rows, cols, Ls, Us,
                                           a + (size_t)(c0 + u0) * lda + c0 + t0, lda);
                else
                    tile_update_valu<T, W>(w, rows, cols, Ls, Us,
                                           a + (size_t)(c0 + u0) * lda + c0 + t0, lda);
                __syncthreads();  // Ls is reloaded by the next tile, Us by the next chunk
            }
        }
    }
}

// X = U^{-1} L^{-1} P B for all nrhs columns of B at once.
//   P: the interchanges composed once into a permutation in LDS (thread 0,
//      pdlapv2.f:208-242's order), then every column gathered through LDS;
//   L (unit lower) top down and U bottom up in 16-row blocks: the diagonal
//      block by substitution in registers, lane = row of one of four
//      columns per wave (four waves: sixteen columns at a time), the block's
//      row of the triangle loaded before the substitution; then every thread
//      updates rows outside the block for one column (dot product over the
//      block).
template <typename T>
__device__ void solve_all(int n, const T *__restrict__ a, int lda, const int *__restrict__ piv,
                          T *__restrict__ B, int ldb, int nrhs, int *s_perm, int *s_piv, T *s_x)
{
    constexpr int SB = 16;  // rows per diagonal block
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & (SB - 1), grp = lane / SB;  // row in the block, column of the wave
    for (int i = tid; i < n; i += BT) {
        s_perm[i] = i;
        s_piv[i] = piv[i] - 1;
    }
    __syncthreads();
    if (tid == 0)
        for (int j = 0; j < n; j++) {
            const int p = s_piv[j];
            if (p != j) {
                const int t = s_perm[j];
                s_perm[j] = s_perm[p];
                s_perm[p] = t;
            }
        }
    __syncthreads();
    for (int c = 0; c < nrhs; c++) {
        T *x = B + (size_t)c * ldb;
        for (int i = tid; i < n; i += BT) s_x[i] = x[s_perm[i]];
        __syncthreads();
        for (int i = tid; i < n; i += BT) x[i] = s_x[i];
        __syncthreads();
    }
    // L, 16-row blocks top down
    for (int b0 = 0; b0 < n; b0 += SB) {
        const int rows = min(SB, n - b0);
        for (int c = 4 * wave + grp; c - grp < nrhs; c += 16) {
            const bool on = c < nrhs && lr < rows;
            T *x = B + (size_t)(on ? c : 0) * ldb;
            const T *ar = a + (size_t)(b0 + lr) + (size_t)b0 * lda;
            T lrow[SB];
#pragma unroll
            for (int i = 0; i < SB; i++) lrow[i] = (on && i < lr) ? ar[(size_t)i * lda] : T(0);
            T v = on ? x[b0 + lr] : T(0);
#pragma unroll
            for (int i = 0; i < SB; i++) {
                const T xi = __shfl(v, i + SB * grp);
                if (lr > i) v = fma(-lrow[i], xi, v);
            }
            if (on) x[b0 + lr] = v;
        }
        __syncthreads();
        const int below = n - b0 - rows;
        for (int idx = tid; idx < below * nrhs; idx += BT) {
            const int i = b0 + rows + idx % below, c = idx / below;
            T *x = B + (size_t)c * ldb;
            T acc = T(0);
            for (int k = 0; k < rows; k++) acc = fma(a[(size_t)i + (size_t)(b0 + k) * lda], x[b0 + k], acc);
            x[i] -= acc;
        }
        __syncthreads();
    }
    // U (non-unit), 16-row blocks bottom up
    const int nb = (n + SB - 1) / SB;
    for (int bb = nb - 1; bb >= 0; bb--) {
        const int b0 = bb * SB, rows = min(SB, n - b0);
        for (int c = 4 * wave + grp; c - grp < nrhs; c += 16) {
            const bool on = c < nrhs && lr < rows;
            T *x = B + (size_t)(on ? c : 0) * ldb;
            const T *ar = a + (size_t)(b0 + lr) + (size_t)b0 * lda;
            T urow[SB];
#pragma unroll
            for (int i = 0; i < SB; i++)
                urow[i] = (on && i > lr && i < rows) ? ar[(size_t)i * lda] : T(0);
            const T d = on ? ar[(size_t)lr * lda] : T(1);
            T v = on ? x[b0 + lr] : T(0);
#pragma unroll
            for (int i = SB - 1; i >= 0; i--) {
                if (lr == i && i < rows) v = v / d;
                const T xi = __shfl(v, i + SB * grp);
                if (lr < i && i < rows) v = fma(-urow[i], xi, v);
            }
            if (on) x[b0 + lr] = v;
        }
        __syncthreads();
        for (int idx = tid; idx < b0 * nrhs; idx += BT) {
            const int i = idx % b0, c = idx / b0;
            T *x = B + (size_t)c * ldb;
            T acc = T(0);
            for (int k = 0; k < rows; k++) acc = fma(a[(size_t)i + (size_t)(b0 + k) * lda], x[b0 + k], acc);
            x[i] -= acc;
        }
        __syncthreads();
    }
}

template <typename T, int W>
__global__ __launch_bounds__(BT) void gesv_batched_kernel(int batch, const int *__restrict__ order,
                                                          const SysDesc *__restrict__ desc,
                                                          T *const *__restrict__ A,
                                                          int *const *__restrict__ ipiv,
                                                          T *const *__restrict__ B,
                                                          int *__restrict__ info, int *queue,
                                                          int flags)
{
    __shared__ T Ls[UT][W + 1];
    __shared__ T Us[W][UT + 1];
    __shared__ T s_v[9];  // [0..3] wave maxima, [4] A(j, j), [5..8] signed values
    __shared__ int s_i[4];
    __shared__ int s_q;
    __shared__ RegPanelLds<T, W> rp;
    __shared__ T L11s[W][W + 1];
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(queue, 1);
        __syncthreads();
        const int q = s_q;
        __syncthreads();  // s_q is rewritten by the next ticket
        if (q >= batch) break;
        const int s = order[q];
        const SysDesc d = desc[s];
        T *a = A[s];
        int *piv = ipiv[s];
        // n < 0 marks a system whose sizes the allocation-free entry rejected
        int inf = d.n < 0 ? -1 : 0;
        if (d.n > 0) factor_system<T, W>(d.n, a, d.lda, piv, inf, Ls, Us, L11s, s_v, s_i, rp, flags);
        if (threadIdx.x == 0) info[s] = inf;
        __syncthreads();
    }
}

// The solves of every factored system (INFO = 0), by the same kind of
// persistent queue (queue[1]), after the factorization launch: a kernel of
// its own keeps its registers out of the factorization's.
template <typename T>
__global__ __launch_bounds__(BT) void getrs_batched_kernel(int batch, const int *__restrict__ order,
                                                           const SysDesc *__restrict__ desc,
                                                           T *const *__restrict__ A,
                                                           int *const *__restrict__ ipiv,
                                                           T *const *__restrict__ B,
                                                           const int *__restrict__ info, int *queue)
{
    __shared__ int s_q;
    __shared__ int s_perm[BATCHED_NMAX], s_piv[BATCHED_NMAX];
    __shared__ T s_x[BATCHED_NMAX];
    for (;;) {
        if (threadIdx.x == 0) s_q = atomicAdd(queue + 1, 1);
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        if (q >= batch) break;
        const int s = order[q];
        const SysDesc d = desc[s];
        if (d.n > 0 && d.nrhs > 0 && info[s] == 0)
            solve_all<T>(d.n, A[s], d.lda, ipiv[s], B[s], d.ldb, d.nrhs, s_perm, s_piv, s_x);
        __syncthreads();
    }
}

__global__ void zero_queue_kernel(int *q)
{
    if (threadIdx.x < 2 && blockIdx.x == 0) q[threadIdx.x] = 0;
}

// Set-up of the allocation-free entry, one workgroup: validates the device
// size arrays (a rejected system gets n = -1, i.e. INFO = -1, and no work),
// orders the systems largest n first by a counting sort (the order only
// balances the queue: every system's result is independent of it) and resets
// the queue.  Buckets: NMAX - n for a valid n, NMAX + 1 for a rejected one.
constexpr int PREP_BUCKETS = BATCHED_NMAX + 2;

__global__ __launch_bounds__(1024) void noalloc_prep_kernel(int batch, const int *__restrict__ n,
                                                            const int *__restrict__ nrhs,
                                                            const int *__restrict__ lda,
                                                            const int *__restrict__ ldb,
                                                            int *queue, int *order, SysDesc *desc)
{
    __shared__ int cnt[PREP_BUCKETS];
    for (int b = threadIdx.x; b < PREP_BUCKETS; b += blockDim.x) cnt[b] = 0;
    if (threadIdx.x < 2) queue[threadIdx.x] = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < batch; s += blockDim.x) {
        const int ns = n[s], rs = nrhs[s], la = lda[s], lb = ldb[s];
        const bool ok = ns >= 0 && ns <= BATCHED_NMAX && rs >= 0 && rs <= BATCHED_NRHS_MAX &&
                        la >= max(1, ns) && lb >= max(1, ns);
        desc[s] = ok ? SysDesc{ns, rs, la, lb} : SysDesc{-1, 0, 1, 1};
        atomicAdd(&cnt[ok ? BATCHED_NMAX - ns : BATCHED_NMAX + 1], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive prefix sum of the bucket counts
        int run = 0;
        for (int b = 0; b < PREP_BUCKETS; b++) {
            const int c = cnt[b];
            cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < batch; s += blockDim.x) {
        const int ns = desc[s].n;  // written by this thread above
        order[atomicAdd(&cnt[ns >= 0 ? BATCHED_NMAX - ns : BATCHED_NMAX + 1], 1)] = s;
    }
}

// ---------------------------------------------------------------------------
// Synthetic systems for the tuner: A uniform[-1, 1) plus n I, B uniform.
__device__ __forceinline__ double hash_uniform(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
template <typename T>
__global__ void fill_synthetic_kernel(int batch, const SysDesc *__restrict__ desc, T *const *A,
                                      T *const *B)
{
    for (int s = blockIdx.y; s < batch; s += gridDim.y) {
        const SysDesc d = desc[s];
        const size_t na = (size_t)d.lda * d.n, nbv = (size_t)d.ldb * d.nrhs;
        for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < na + nbv;
             e += (size_t)gridDim.x * 256) {
            const double u = hash_uniform(((uint64_t)s << 32) ^ e);
            if (e < na) {
                const int i = (int)(e % d.lda), j = (int)(e / d.lda);
                A[s][e] = (T)(u + (i == j ? (double)d.n : 0.0));
            } else {
                B[s][e - na] = (T)u;
            }
        }
    }
}

}  // namespace

// ============================================================================
// Host side
// ============================================================================
struct BatchedPlan {
    int batch = 0;
    std::vector<int> n, nrhs, lda, ldb;
    int *d_order = nullptr;
    SysDesc *d_desc = nullptr;
    int *d_queue = nullptr;
    void *graph_exec = nullptr;  // hipGraphExec_t of the last captured launch
    const void *g_key[4] = {nullptr, nullptr, nullptr, nullptr};
    int g_prec = -1, g_w = 0, g_wgs = 0;
    ~BatchedPlan()
    {
        if (graph_exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec));
        if (d_order) (void)hipFree(d_order);
        if (d_desc) (void)hipFree(d_desc);
        if (d_queue) (void)hipFree(d_queue);
    }
};

namespace {

bool same_sizes(const BatchedPlan &p, int batch, const int *n, const int *nrhs, const int *lda,
                const int *ldb)
{
    if (p.batch != batch) return false;
    for (int s = 0; s < batch; s++)
        if (p.n[s] != n[s] || p.nrhs[s] != nrhs[s] || p.lda[s] != lda[s] || p.ldb[s] != ldb[s])
            return false;
    return true;
}

template <typename T>
void launch(const BatchedPlan &p, int w, int wgs, int solve_wgs, T *const *A, int *const *ipiv,
            T *const *B, int *info, hipStream_t s)
{
    hipLaunchKernelGGL(zero_queue_kernel, dim3(1), dim3(64), 0, s, p.d_queue);
#define SLA_BATCHED(WV)                                                                          \
    hipLaunchKernelGGL((gesv_batched_kernel<T, WV>), dim3(wgs), dim3(BT), 0, s, p.batch,         \
                       p.d_order, p.d_desc, A, ipiv, B, info, p.d_queue, batched_flags())
    switch (w) {
    case 8: SLA_BATCHED(8); break;
    case 16: SLA_BATCHED(16); break;
    case 32: SLA_BATCHED(32); break;
    default: throw std::invalid_argument("batched: panel width must be 8, 16 or 32");
    }
#undef SLA_BATCHED
    SLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(getrs_batched_kernel<T>, dim3(solve_wgs), dim3(BT), 0, s, p.batch,
                       p.d_order, p.d_desc, (T *const *)A, ipiv, B, info, p.d_queue);
    SLA_LAUNCH_CHECK();
}

}  // namespace

BatchedPlan &batched_plan(Handle &h, int batch, const int *n, const int *nrhs, const int *lda,
                          const int *ldb)
{
    if (!h.batched) h.batched = new BatchedPlan();
    BatchedPlan &p = *h.batched;
    if (same_sizes(p, batch, n, nrhs, lda, ldb)) return p;
    for (int s = 0; s < batch; s++) {
        if (n[s] < 0 || n[s] > BATCHED_NMAX || nrhs[s] < 0 || nrhs[s] > BATCHED_NRHS_MAX ||
            lda[s] < std::max(1, n[s]) || ldb[s] < std::max(1, n[s]))
            throw std::invalid_argument("batched: system " + std::to_string(s) +
                                        " has n, nrhs, lda or ldb out of range");
    }
    SLA_HIP(hipStreamSynchronize(h.stream));
    if (p.graph_exec) {
        SLA_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(p.graph_exec)));
        p.graph_exec = nullptr;
    }
    if (p.d_order) SLA_HIP(hipFree(p.d_order));
    if (p.d_desc) SLA_HIP(hipFree(p.d_desc));
    if (!p.d_queue) SLA_HIP(hipMalloc(&p.d_queue, 64));
    p.batch = batch;
    p.n.assign(n, n + batch);
    p.nrhs.assign(nrhs, nrhs + batch);
    p.lda.assign(lda, lda + batch);
    p.ldb.assign(ldb, ldb + batch);
    // largest work first (n^3 + n^2 nrhs), ties by index
    std::vector<int> order(batch);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const double wx = (double)n[x] * n[x] * (n[x] + 3.0 * nrhs[x]);
        const double wy = (double)n[y] * n[y] * (n[y] + 3.0 * nrhs[y]);
        return wx > wy;
    });
    std::vector<SysDesc> desc(batch);
    for (int s = 0; s < batch; s++) desc[s] = SysDesc{n[s], nrhs[s], lda[s], ldb[s]};
    SLA_HIP(hipMalloc(&p.d_order, sizeof(int) * std::max(1, batch)));
    SLA_HIP(hipMalloc(&p.d_desc, sizeof(SysDesc) * std::max(1, batch)));
    SLA_HIP(hipMemcpy(p.d_order, order.data(), sizeof(int) * batch, hipMemcpyHostToDevice));
    SLA_HIP(hipMemcpy(p.d_desc, desc.data(), sizeof(SysDesc) * batch, hipMemcpyHostToDevice));
    return p;
}

void batched_plan_free(BatchedPlan *p) { delete p; }

}  // namespace sla

// ---------------------------------------------------------------------------
// Allocation-free entry: the queue counter, the order and the size
// descriptors live in this fixed device array (one copy per GPU, kept until
// the process ends) instead of memory the library allocates.  Layout: queue
// at 0, order at 256, descriptors after BATCHED_NOALLOC_MAX ints.
static __device__ __align__(256) unsigned char
    scratch[256 + (size_t)sla::BATCHED_NOALLOC_MAX * (sizeof(int) + 4 * sizeof(int))];

namespace sla {

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};  // device address of `scratch` on each GPU

// Device address of this GPU's `scratch` (looked up once per device, with
// that device current; host code never uses the array's name as an address).
void *noalloc_scratch(int device)
{
    if (device < 0 || device >= 64) throw std::runtime_error("batched: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        SLA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("batched: no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace

template <typename T>
void gesv_batched_noalloc_device(Handle &h, int batch, const int *dn, const int *dnrhs,
                                 T *const *dA, const int *dlda, int *const *dipiv, T *const *dB,
                                 const int *dldb, int *dinfo)
{
    if (batch < 0 || batch > BATCHED_NOALLOC_MAX)
        throw std::invalid_argument("batched (no allocation): batch above " +
                                    std::to_string(BATCHED_NOALLOC_MAX) + " systems");
    int cur = -1;
    SLA_HIP(hipGetDevice(&cur));
    if (cur != h.device)
        throw std::runtime_error("batched (no allocation): the handle's device must be current");
    if (batch == 0) return;
    static_assert(sizeof(SysDesc) == 4 * sizeof(int), "descriptor layout of the scratch array");
    char *base = static_cast<char *>(noalloc_scratch(cur));
    int *queue = reinterpret_cast<int *>(base);
    int *order = reinterpret_cast<int *>(base + 256);
    SysDesc *desc = reinterpret_cast<SysDesc *>(base + 256 + sizeof(int) * BATCHED_NOALLOC_MAX);
    hipLaunchKernelGGL(noalloc_prep_kernel, dim3(1), dim3(1024), 0, h.stream, batch, dn, dnrhs,
                       dlda, dldb, queue, order, desc);
    SLA_LAUNCH_CHECK();
    const int wgs = std::max(1, h.batched_wgs_per_cu) * h.num_cus;
#define SLA_BATCHED(WV)                                                                          \
    hipLaunchKernelGGL((gesv_batched_kernel<T, WV>), dim3(wgs), dim3(BT), 0, h.stream, batch,     \
                       order, desc, dA, dipiv, dB, dinfo, queue, batched_flags())
    switch (h.batched_w) {
    case 8: SLA_BATCHED(8); break;
    case 16: SLA_BATCHED(16); break;
    case 32: SLA_BATCHED(32); break;
    default: throw std::invalid_argument("batched: panel width must be 8, 16 or 32");
    }
#undef SLA_BATCHED
    SLA_LAUNCH_CHECK();
    hipLaunchKernelGGL(getrs_batched_kernel<T>, dim3(4 * h.num_cus), dim3(BT), 0, h.stream, batch,
                       order, desc, (T *const *)dA, (int *const *)dipiv, dB, dinfo, queue);
    SLA_LAUNCH_CHECK();
}

template void gesv_batched_noalloc_device<double>(Handle &, int, const int *, const int *,
                                                  double *const *, const int *, int *const *,
                                                  double *const *, const int *, int *);
template void gesv_batched_noalloc_device<float>(Handle &, int, const int *, const int *,
                                                 float *const *, const int *, int *const *,
                                                 float *const *, const int *, int *);

template <typename T>
void gesv_batched_device(Handle &h, int batch, const int *n, const int *nrhs, T *const *dA,
                         const int *lda, int *const *dipiv, T *const *dB, const int *ldb, int *dinfo)
{
    if (batch <= 0) return;
    BatchedPlan &p = batched_plan(h, batch, n, nrhs, lda, ldb);
    const int w = h.batched_w;
    const int wgs = std::max(1, h.batched_wgs_per_cu) * h.num_cus;
    const int prec = sizeof(T) == 8 ? 0 : 1;
    if (!h.batched_graph) {
        launch<T>(p, w, wgs, 4 * h.num_cus, dA, dipiv, dB, dinfo, h.stream);
        SLA_HIP(hipStreamSynchronize(h.stream));
        return;
    }
    // hipGraph path: the queue reset (a kernel, not a memset) and the
    // persistent launch captured once per (pointer arrays, sizes, shape)
    // and replayed afterwards.
    const void *key[4] = {dA, dipiv, dB, dinfo};
    const bool hit = p.graph_exec && std::equal(key, key + 4, p.g_key) && p.g_prec == prec &&
                     p.g_w == w && p.g_wgs == wgs;
    if (!hit) {
        if (p.graph_exec) {
            SLA_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(p.graph_exec)));
            p.graph_exec = nullptr;
        }
        hipStream_t cs = h.own_stream;  // capture on the handle's own stream
        SLA_HIP(hipStreamSynchronize(h.stream));
        hipGraph_t g = nullptr;
        SLA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        try {
            launch<T>(p, w, wgs, 4 * h.num_cus, dA, dipiv, dB, dinfo, cs);
        } catch (...) {
            (void)hipStreamEndCapture(cs, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SLA_HIP(hipStreamEndCapture(cs, &g));
        hipGraphExec_t ge = nullptr;
        SLA_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        SLA_HIP(hipGraphDestroy(g));
        p.graph_exec = ge;
        std::copy(key, key + 4, p.g_key);
        p.g_prec = prec;
        p.g_w = w;
        p.g_wgs = wgs;
    }
    SLA_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(p.graph_exec), h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
}

template void gesv_batched_device<double>(Handle &, int, const int *, const int *, double *const *,
                                          const int *, int *const *, double *const *, const int *,
                                          int *);
template void gesv_batched_device<float>(Handle &, int, const int *, const int *, float *const *,
                                         const int *, int *const *, float *const *, const int *,
                                         int *);

// Autotuner: times every (panel width, workgroups per CU) on synthetic
// systems of the given sizes (diagonally dominant, so every candidate does
// the same work) and keeps the fastest in the handle.
template <typename T>
BatchedTune tune_batched_device(Handle &h, int batch, const int *n, const int *nrhs, int reps)
{
    BatchedTune out;
    if (batch <= 0) return out;
    std::vector<int> lda(batch), ldb(batch);
    size_t total = 0, total_piv = 0;
    for (int s = 0; s < batch; s++) {
        lda[s] = std::max(1, n[s]);
        ldb[s] = std::max(1, n[s]);
        total += (size_t)lda[s] * n[s] + (size_t)ldb[s] * nrhs[s];
        total_piv += (size_t)std::max(1, n[s]);
    }
    BatchedPlan &p = batched_plan(h, batch, n, nrhs, lda.data(), ldb.data());
    T *data = nullptr;
    int *piv = nullptr, *info = nullptr;
    void **ptrs = nullptr;
    SLA_HIP(hipMalloc(&data, total * sizeof(T) + 256));
    SLA_HIP(hipMalloc(&piv, total_piv * sizeof(int) + 256));
    SLA_HIP(hipMalloc(&info, sizeof(int) * batch));
    SLA_HIP(hipMalloc(&ptrs, sizeof(void *) * 3 * batch));
    std::vector<void *> hp(3 * (size_t)batch);
    size_t off = 0, offp = 0;
    for (int s = 0; s < batch; s++) {
        hp[s] = data + off;
        off += (size_t)lda[s] * n[s];
        hp[2 * (size_t)batch + s] = data + off;
        off += (size_t)ldb[s] * nrhs[s];
        hp[(size_t)batch + s] = piv + offp;
        offp += (size_t)std::max(1, n[s]);
    }
    SLA_HIP(hipMemcpy(ptrs, hp.data(), sizeof(void *) * hp.size(), hipMemcpyHostToDevice));
    T *const *dA = reinterpret_cast<T *const *>(ptrs);
    int *const *dP = reinterpret_cast<int *const *>(ptrs + batch);
    T *const *dB = reinterpret_cast<T *const *>(ptrs + 2 * (size_t)batch);
    hipEvent_t e0, e1;
    SLA_HIP(hipEventCreate(&e0));
    SLA_HIP(hipEventCreate(&e1));
    double best = 1e30;
    for (int w : {8, 16, 32})
        for (int k : {1, 2, 4}) {
            double ms_sum = 0.0;
            for (int r = 0; r < reps + 1; r++) {
                hipLaunchKernelGGL(fill_synthetic_kernel<T>, dim3(64, std::min(batch, 65535)), dim3(256), 0, h.stream,
                                   batch, p.d_desc, dA, dB);
                SLA_HIP(hipEventRecord(e0, h.stream));
                launch<T>(p, w, k * h.num_cus, 4 * h.num_cus, dA, dP, dB, info, h.stream);
                SLA_HIP(hipEventRecord(e1, h.stream));
                SLA_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                SLA_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) ms_sum += ms;  // the first run warms up
            }
            const double ms = ms_sum / std::max(1, reps);
            out.tried.push_back({w, k, ms});
            if (ms < best) {
                best = ms;
                out.w = w;
                out.wgs_per_cu = k;
                out.ms = ms;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(data);
    (void)hipFree(piv);
    (void)hipFree(info);
    (void)hipFree(ptrs);
    h.batched_w = out.w;
    h.batched_wgs_per_cu = out.wgs_per_cu;
    return out;
}

template BatchedTune tune_batched_device<double>(Handle &, int, const int *, const int *, int);
template BatchedTune tune_batched_device<float>(Handle &, int, const int *, const int *, int);

}  // namespace sla
