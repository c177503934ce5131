// panel_tslu.hip -- the panel leaf of the low-precision factorization with
// tournament pivoting (communication-avoiding LU of a tall leaf): the W pivot
// rows of a leaf are chosen by a reduction tree of small partial-pivoting
// eliminations instead of W device-wide pivot searches, so a leaf costs one
// device-wide barrier instead of one per column.
//
// The FP64 path keeps PDGETF2's partial pivoting (panel.hip), whose pivots
// must equal the reference's.  This leaf serves the mixed-precision solve,
// where only the backward error of the refined result is prescribed; the
// factors it produces are an exact LU (P A = L U up to rounding, |L| <= 1 not
// guaranteed but in practice bounded) of the same kind partial pivoting gives.
//
// One launch of G <= 32 workgroups of 512 threads; the leaf's m rows are
// spread as in panel.hip (row g*R + q*512 + tid, R = RPT*512).
//   1. every workgroup runs partial-pivoting elimination on its own rows in
//      registers (per column: wave argmax, two workgroup barriers, the pivot
//      row through LDS) and so chooses W candidate rows, published with their
//      original values; workgroup 0 also publishes the leaf's top W rows;
//   2. one device-wide barrier;
//   3. every workgroup runs the same elimination over all G*W candidates and
//      so agrees on the W pivot rows and on their factors L11\U11 (the
//      elimination of the pivot rows, in pivot order) without another barrier;
//   4. the pivots become W sequential interchanges (IPIV as PDGETRF returns
//      it); every workgroup writes its rows: the top W rows are L11\U11, every
//      other row r is L(r) = A(src(r)) U11^{-1}, where src(r) is the row the
//      interchanges bring to position r (a displaced top row comes from the
//      published copy, any other row is the workgroup's own); the
//      interchanges reach the rest of the block column as their net moves
//      (at most 2W rows change place), one gather and one scatter a column.
#include "sla_common.h"
#include "sla_kernels.h"

#include <climits>

// wavefront-wide maximum of the ROCm device library (declared by the HIP
// headers only with their optional 64-bit warp-sync types)
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(
    unsigned long long);

namespace sla {

namespace {

constexpr int TPT = 512;         // threads per workgroup
constexpr int TNW = TPT / 64;    // waves per workgroup
constexpr int TW = 32;           // leaf width
constexpr int TMR = 2;           // candidate rows per thread in the tournament
constexpr int TGMAX = TPT * TMR / TW;  // workgroups per launch at most (32)
// candidate values are stored column by column (value c of candidate sid at
// c * TCAND + sid), so the tournament's loads are contiguous across lanes
constexpr int TCAND = TGMAX * TW;

// Pivot key of a row: |v| as float bits (monotone for |v| >= 0; a NaN counts
// as zero) in the upper half, INT_MAX - row in the lower, so the larger key
// is the larger magnitude and, on a tie, the lower row; 0 = no row.
__device__ __forceinline__ uint64_t piv_key(float v, int row)
{
    const float a = fabsf(v);
    const uint32_t ab = (a == a) ? __float_as_uint(a) : 0u;
    return row == INT_MAX ? 0ull : (((uint64_t)ab << 32) | (uint32_t)(INT_MAX - row));
}

// Partial-pivoting elimination over the rows a whole workgroup holds (NR
// per thread; key = row identity and tie-break, INT_MAX = no row; pid =
// payload id).  Selects up to wa pivots in order: out_key[j], out_pid[j]
// (written by thread 0); with s_u, row j of U (columns j..) goes to s_u[j].
// Per column: one wave max over the packed pivot keys, the winning lane of
// every wave posts its row to LDS, one workgroup barrier, every thread folds
// the 8 posts in the same order and applies the rank-1 update with the
// winner's row.  The posts are double-buffered by column parity: a buffer is
// rewritten two columns later, after a barrier every reader has passed.
//
// The columns rotate through the registers: at step j, register column 0
// holds matrix column j, and the update writes column c+1's new value into
// register column c, so every register index is a compile-time constant while
// the step loop stays rolled.
struct alignas(16) GeppLds {
    uint64_t wkey[2][TNW];
    float wrow[2][TNW][TW];
    int wpid[2][TNW];
};
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef unsigned long long u2_t __attribute__((ext_vector_type(2)));

template <typename T, int NR, bool FULL>
__device__ __forceinline__ void wg_gepp(T (&v)[NR][TW], const int (&key)[NR], const int (&pid)[NR],
                                        int wa, int *out_key, int *out_pid, GeppLds &sl,
                                        T (*s_u)[TW + 1])
{
    static_assert(sizeof(T) == 4, "the packed pivot key holds a float magnitude");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    bool act[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) act[q] = key[q] != INT_MAX;
    const int steps = FULL ? TW : wa;
#pragma unroll 1
    for (int j = 0; j < steps; j++) {
        const int pb = j & 1;
        uint64_t bk = 0;
        int bq = 0;
#pragma unroll
        for (int q = 0; q < NR; q++) {
            const uint64_t k = act[q] ? piv_key(v[q][0], key[q]) : 0ull;
            if (k > bk) {
                bk = k;
                bq = q;
            }
        }
        const uint64_t wmax = __ockl_wfred_max_u64(bk);
        if (wmax != 0 && bk == wmax) {  // the one lane holding the wave's winner
#pragma unroll
            for (int c4 = 0; c4 < TW; c4 += 4) {
                f4_t sel4;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    T sel = v[0][c4 + e];
#pragma unroll
                    for (int q = 1; q < NR; q++) sel = (bq == q) ? v[q][c4 + e] : sel;
                    sel4[e] = sel;
                }
                *reinterpret_cast<f4_t *>(&sl.wrow[pb][wave][c4]) = sel4;
            }
            int sp = pid[0];
#pragma unroll
            for (int q = 1; q < NR; q++) sp = (bq == q) ? pid[q] : sp;
            sl.wpid[pb][wave] = sp;
        }
        if (lane == 0) sl.wkey[pb][wave] = wmax;
        __syncthreads();
        u2_t kk[TNW / 2];
#pragma unroll
        for (int w = 0; w < TNW / 2; w++) kk[w] = reinterpret_cast<const u2_t *>(sl.wkey[pb])[w];
        uint64_t fk = kk[0][0];
        int fw = 0;
#pragma unroll
        for (int w = 1; w < TNW; w++) {
            const uint64_t k = kk[w / 2][w % 2];
            if (k > fk) {
                fk = k;
                fw = w;
            }
        }
        // fk == 0 (uniform): no rows left -- no pivot, and every multiplier
        // below is 0 because no row is active
        const bool any = fk != 0;
        if (any && bk == fk) {  // the owner of the pivot row retires it
#pragma unroll
            for (int q = 0; q < NR; q++)
                if (bq == q) act[q] = false;
        }
        const float *prow = sl.wrow[pb][fw];
        if (tid == 0) {
            out_key[j] = any ? INT_MAX - (int)(uint32_t)fk : INT_MAX;
            out_pid[j] = any ? sl.wpid[pb][fw] : -1;
        }
        if (any && s_u != nullptr && tid < TW - j) s_u[j][j + tid] = prow[tid];
        T pr[TW];
#pragma unroll
        for (int c4 = 0; c4 < TW; c4 += 4) {
            const f4_t x = reinterpret_cast<const f4_t *>(prow)[c4 / 4];
#pragma unroll
            for (int e = 0; e < 4; e++) pr[c4 + e] = x[e];
        }
        const T pj = pr[0];
        const bool nz = pj != T(0);
        const T rcp = nz ? T(1) / pj : T(0);
        // every row rotates (a retired row with multiplier 0: its values are
        // never read again), so the registers need no copies between steps
#pragma unroll
        for (int q = 0; q < NR; q++) {
            T l = v[q][0];
            if (nz) l = l * rcp;
            l = act[q] ? l : T(0);
#pragma unroll
            for (int c = 1; c < TW; c++) v[q][c - 1] = fma(-l, pr[c], v[q][c]);
            v[q][TW - 1] = T(0);
        }
    }
    __syncthreads();  // out_* and s_u complete before anyone reads them
}

template <bool FULL>
__device__ __forceinline__ int colx(int c, int wa)
{
    return FULL ? c : min(c, wa - 1);
}

// FULL: the leaf is TW columns wide (no column predicates).
template <typename T, int RPT, bool FULL>
__global__ __launch_bounds__(TPT) void panel_tslu_kernel(
    T *__restrict__ A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
    int *__restrict__ ipiv, TsluSlots<T> slots, unsigned *ctr, unsigned target, int *abort_flag,
    int *zero_col, uint64_t spin_limit)
{
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int R = RPT * TPT;
    if (FULL) wa = TW;
    // column c of the leaf as loaded is colx<FULL>(c, wa): columns right of a
    // narrow leaf repeat its last column (never written back, never pivoted on)

    __shared__ GeppLds s_gl;
    __shared__ int s_k2[TW], s_p2[TW];
    __shared__ T s_lu[TW][TW + 1];
    __shared__ int s_dst[2 * TW], s_src[2 * TW], s_np;

    if (!FULL)
        for (int idx = tid; idx < TW * (TW + 1); idx += TPT) s_lu[idx / (TW + 1)][idx % (TW + 1)] = T(0);
    if (tid < TW) s_k2[tid] = INT_MAX;
    __syncthreads();

    // 1. elimination of the workgroup's own rows: its TW candidates
    {
        T v[RPT][TW];
        int key[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * TPT + tid;
            key[q] = r < m ? r : INT_MAX;
            const int rr = r < m ? r : m - 1;
#pragma unroll
            for (int c = 0; c < TW; c++) v[q][c] = A[rr + (size_t)colx<FULL>(c, wa) * lda];
        }
        wg_gepp<T, RPT, FULL>(v, key, key, wa, s_k2, s_p2, s_gl, nullptr);
    }
    // publish the workgroup's candidates (ids, original values); workgroup 0
    // also the leaf's top rows
    for (int idx = tid; idx < TW * TW; idx += TPT) {
        const int c = idx / TW, j = idx % TW;  // consecutive threads: consecutive candidates
        const int r = (j < wa) ? s_k2[j] : INT_MAX;
        const T x = r != INT_MAX ? A[r + (size_t)colx<FULL>(c, wa) * lda] : T(0);
        coherent_store(&slots.cand_v[(size_t)c * TCAND + g * TW + j], x);
        if (c == 0) coherent_store(&slots.cand_r[g * TW + j], r);
        if (g == 0) {
            const T y = (j < wa) ? A[j + (size_t)colx<FULL>(c, wa) * lda] : T(0);
            coherent_store(&slots.top_v[j * TW + c], y);
        }
    }

    // 2. device-wide barrier
    if (!grid_sync(ctr, target, abort_flag, spin_limit)) return;

    // 3. the tournament, identical in every workgroup: one elimination over
    //    all G*TW candidates (original values) fixes the pivot rows, their
    //    order and U11 (the pivot rows as the elimination leaves them); L11
    //    is then solved from the pivot rows' original values
    if (tid < TW) s_k2[tid] = INT_MAX;
    __syncthreads();
    {
        const int total = G * TW;
        T v[TMR][TW];
        int key[TMR], pid[TMR];
#pragma unroll
        for (int q = 0; q < TMR; q++) {
            const int sid = q * TPT + tid;
            const bool ok = sid < total;
            const int r = ok ? coherent_load(&slots.cand_r[sid]) : INT_MAX;
            key[q] = r;
            pid[q] = ok ? sid : -1;
            const int sv = (r != INT_MAX) ? sid : 0;
#pragma unroll
            for (int c = 0; c < TW; c++) v[q][c] = coherent_load(&slots.cand_v[(size_t)c * TCAND + sv]);
        }
        wg_gepp<T, TMR, FULL>(v, key, pid, wa, s_k2, s_p2, s_gl, s_lu);
    }
    if (wave == 0 && lane < wa) {
        // L11: lane r solves its pivot row's original values against U11
        // (x U11^{-1}; the first r entries are row r's multipliers), the
        // same computation every other row of the leaf gets below
        T x[TW];
        const int sid = s_p2[lane];
#pragma unroll
        for (int c = 0; c < TW; c++) x[c] = coherent_load(&slots.cand_v[(size_t)c * TCAND + sid]);
#pragma unroll
        for (int c = 0; c < TW; c++) {
            if (c < lane) {
                T acc = x[c];
#pragma unroll
                for (int k = 0; k < c; k++) acc = fma(-x[k], s_lu[k][c], acc);
                const T u = s_lu[c][c];
                x[c] = (u != T(0)) ? acc / u : acc;
                s_lu[lane][c] = x[c];
            }
        }
    }
    __syncthreads();

    // 5a. the pivots as sequential interchanges (wave 0): lane l < wa holds
    //     position l, lane wa + k the position of pivot k when it lies below
    //     the top rows; `at` is the row currently at that position
    if (wave == 0) {
        int pos = -1;
        if (lane < wa) {
            pos = lane;
        } else if (lane < 2 * wa) {
            const int pk = s_k2[lane - wa];
            pos = (pk >= wa) ? pk : -1;
        }
        int at = pos;
        for (int k = 0; k < wa; k++) {
            const int pk = s_k2[k];
            const unsigned long long bx = __ballot(pos >= 0 && at == pk);
            const int xl = bx ? __builtin_ctzll(bx) : k;  // always found: pivots are distinct rows
            const int xpos = __builtin_amdgcn_readlane(pos, xl);
            if (lane == 0 && g == 0) ipiv[row0 + k] = row0 + xpos + 1;
            const int at_k = __builtin_amdgcn_readlane(at, k);
            const int at_x = __builtin_amdgcn_readlane(at, xl);
            if (lane == k) at = at_x;
            if (lane == xl) at = at_k;
        }
        const bool moved = pos >= 0 && at != pos;
        const unsigned long long bm = __ballot(moved);
        const int slot = __builtin_popcountll(bm & ((1ull << lane) - 1ull));
        if (moved) {
            s_dst[slot] = pos;
            s_src[slot] = at;
        }
        if (lane == 0) s_np = __builtin_popcountll(bm);
        if (g == 0 && lane < wa && s_lu[lane][lane] == T(0)) atomicMin(zero_col, col0 + lane + 1);
    }
    __syncthreads();

    // 5b. the workgroup's rows of the leaf
    const int np = s_np;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = g * R + q * TPT + tid;
        if (r < m) {
            T x[TW];
            if (r < wa) {
#pragma unroll
                for (int c = 0; c < TW; c++) x[c] = s_lu[r][c];
            } else {
                int src = r;
                for (int i = 0; i < np; i++)
                    if (s_dst[i] == r) src = s_src[i];
                if (src == r) {
#pragma unroll
                    for (int c = 0; c < TW; c++) x[c] = A[r + (size_t)colx<FULL>(c, wa) * lda];
                } else {
#pragma unroll
                    for (int c = 0; c < TW; c++) x[c] = coherent_load(&slots.top_v[src * TW + c]);
                }
                // L(r) = x U11^{-1}
                // (columns right of a narrow leaf compute throwaway values)
#pragma unroll
                for (int c = 0; c < TW; c++) {
                    T acc = x[c];
#pragma unroll
                    for (int k = 0; k < c; k++) acc = fma(-x[k], s_lu[k][c], acc);
                    const T u = s_lu[c][c];
                    x[c] = (u != T(0)) ? acc / u : acc;
                }
            }
            if (FULL) {
#pragma unroll
                for (int c = 0; c < TW; c++) A[r + (size_t)c * lda] = x[c];
            } else {
#pragma unroll
                for (int c = 0; c < TW; c++)
                    if (c < wa) A[r + (size_t)c * lda] = x[c];
            }
        }
    }

    // 5c. the interchanges on the block column's other columns, as their net
    //     moves (position s_dst[i] receives the row at s_src[i], at most 2 W
    //     of them): one wave per column, lane i moves one element; the whole
    //     wave's load completes before its store, so no move reads a value
    //     another one already overwrote
    const int ne = lcols + rcols;
    for (int e = g * TNW + wave; e < ne; e += G * TNW) {
        if (lane < np) {
            const int cc = (e < lcols) ? (e - lcols) : (wa + e - lcols);
            T *col = A + (ptrdiff_t)cc * lda;
            const T x = col[s_src[lane]];
            col[s_dst[lane]] = x;
        }
    }
}

}  // namespace

template <typename T>
bool panel_tslu_fits(int m, int wa, int target_wgs)
{
    (void)target_wgs;
    return wa >= 1 && wa <= TW && m >= wa && ceil_div(m, 2 * TPT) <= TGMAX;
}

template <typename T>
unsigned panel_tslu(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                    int *ipiv, const PanelWork<T> &pw, const TsluSlots<T> &slots, unsigned base,
                    hipStream_t s)
{
    if (!panel_tslu_fits<T>(m, wa, pw.target_wgs))
        throw std::logic_error("panel_tslu: leaf shape outside the tournament kernel's range");
    int rpt = 1, G = ceil_div(m, TPT);
    while ((G > pw.target_wgs || G > TGMAX) && rpt < 2) {
        rpt *= 2;
        G = ceil_div(m, TPT * rpt);
    }
    const unsigned target = base + (unsigned)G;
#define SLA_TSLU(RV, FV)                                                                      \
    hipLaunchKernelGGL((panel_tslu_kernel<T, RV, FV>), dim3(G), dim3(TPT), 0, s, A, lda, m, wa, \
                       lcols, rcols, row0, col0, ipiv, slots, pw.ctr, target, pw.abort_flag,   \
                       pw.zero_col, pw.spin_limit)
    if (wa == TW) {
        if (rpt == 1) SLA_TSLU(1, true);
        else SLA_TSLU(2, true);
    } else {
        if (rpt == 1) SLA_TSLU(1, false);
        else SLA_TSLU(2, false);
    }
#undef SLA_TSLU
    SLA_LAUNCH_CHECK();
    return (unsigned)G;
}

template bool panel_tslu_fits<float>(int, int, int);
template unsigned panel_tslu<float>(float *, int, int, int, int, int, int, int, int *,
                                    const PanelWork<float> &, const TsluSlots<float> &, unsigned,
                                    hipStream_t);

}  // namespace sla
