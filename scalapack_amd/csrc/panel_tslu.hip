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
// One launch of G workgroups of 512 threads (8 waves); the leaf's m rows are
// spread as in panel.hip (row g*R + q*512 + tid, R = RPT*512).
//   1. every wave runs partial-pivoting elimination on its own 64*RPT rows,
//      entirely in registers (wave shuffles for the argmax, lane reads for the
//      pivot row), and keeps the W rows it chose;
//   2. wave 0 repeats that on the 8*W rows its workgroup chose (original
//      values), giving W candidates per workgroup, published with their
//      original values; workgroup 0 also publishes the leaf's top W rows;
//   3. one device-wide barrier;
//   4. every workgroup runs the same two-level tournament on the G*W
//      candidates (8 waves on slices, then wave 0 on the 8*W survivors) and
//      so agrees on the W pivot rows and on their factors L11\U11 (the final
//      elimination of the pivot rows, in pivot order) without another barrier;
//   5. the pivots become W sequential interchanges (IPIV as PDGETRF returns
//      it); every workgroup writes its rows: the top W rows are L11\U11, every
//      other row r is L(r) = A(src(r)) U11^{-1}, where src(r) is the row the
//      interchanges bring to position r (a displaced top row comes from the
//      published copy, any other row is the workgroup's own); the
//      interchanges are applied to the rest of the block column in order.
#include "sla_common.h"
#include "sla_kernels.h"

#include <climits>

namespace sla {

namespace {

constexpr int TPT = 512;         // threads per workgroup
constexpr int TNW = TPT / 64;    // waves per workgroup
constexpr int TW = 32;           // leaf width
constexpr int TMR = 2;           // candidate rows per lane in the first tournament level
constexpr int TGMAX = TNW * 64 * TMR / TW;  // workgroups per launch at most (32)

// Partial-pivoting elimination of the NR x W rows a wave holds (NR per lane,
// key = row identity and tie-break, INT_MAX = no row; pid = payload id).
// Selects up to wa pivots in order: out_key[j], out_pid[j] (uniform, written
// by lane 0), and with OUT_LU the pivot row's values when chosen
// (multipliers left of j, U from j on).
//
// The columns rotate through the registers: at step j, register column 0
// holds matrix column j, and the rank-1 update writes column c+1's new value
// into register column c.  Every register index is a compile-time constant
// while the step loop itself stays rolled (one copy of the body).  With
// OUT_LU the multipliers rotate in from the right (mul[.][W-1] is the newest).
template <typename T, int NR, bool OUT_LU, bool FULL>
__device__ __forceinline__ void wave_gepp(T (&v)[NR][TW], const int (&key)[NR],
                                          const int (&pid)[NR], int wa, int *out_key,
                                          int *out_pid, T (*out_lu)[TW + 1])
{
    const int lane = threadIdx.x & 63;
    bool act[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) act[q] = key[q] != INT_MAX;
    T mul[OUT_LU ? NR : 1][TW];
    if constexpr (OUT_LU) {
#pragma unroll
        for (int q = 0; q < NR; q++)
#pragma unroll
            for (int c = 0; c < TW; c++) mul[q][c] = T(0);
    }
    const int steps = FULL ? TW : wa;
#pragma unroll 1
    for (int j = 0; j < steps; j++) {
        double bv = -1.0;
        int bk = INT_MAX, bq = 0;
#pragma unroll
        for (int q = 0; q < NR; q++) {
            if (act[q]) {
                const double x = fabs((double)v[q][0]);
                if (pivot_better(x, key[q], bv, bk)) {
                    bv = x;
                    bk = key[q];
                    bq = q;
                }
            }
        }
        double wv = bv;
        int wk = bk;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(wv, off);
            const int ok = __shfl_xor(wk, off);
            if (pivot_better(ov, ok, wv, wk)) {
                wv = ov;
                wk = ok;
            }
        }
        if (wk == INT_MAX) {  // uniform: no rows left
            if (lane == 0) {
                out_key[j] = INT_MAX;
                out_pid[j] = -1;
            }
            continue;
        }
        const bool mine = (bk == wk);
        const int wl = __builtin_ctzll(__ballot(mine));
        T prow[TW];
#pragma unroll
        for (int c = 0; c < TW; c++) {
            T sel = v[0][c];
#pragma unroll
            for (int q = 1; q < NR; q++) sel = (bq == q) ? v[q][c] : sel;
            prow[c] = lane_read(sel, wl);
        }
        int sp = pid[0];
#pragma unroll
        for (int q = 1; q < NR; q++) sp = (bq == q) ? pid[q] : sp;
        const int wpid = __builtin_amdgcn_readlane(sp, wl);
        if constexpr (OUT_LU) {
            // row j of L11\U11: multipliers l_0..l_{j-1} are mul[TW-j..TW-1]
            // of the pivot row, U(j, j..) is prow[0..TW-1-j]
            T pm[TW];
#pragma unroll
            for (int c = 0; c < TW; c++) {
                T sel = mul[0][c];
#pragma unroll
                for (int q = 1; q < NR; q++) sel = (bq == q) ? mul[q][c] : sel;
                pm[c] = lane_read(sel, wl);
            }
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < TW; c++) {
                    if (c < TW - j) out_lu[j][j + c] = prow[c];
                    if (c >= TW - j) out_lu[j][c - (TW - j)] = pm[c];
                }
            }
        }
        if (lane == 0) {
            out_key[j] = wk;
            out_pid[j] = wpid;
        }
#pragma unroll
        for (int q = 0; q < NR; q++)
            if (mine && bq == q) act[q] = false;
        const T pj = prow[0];
        const bool nz = pj != T(0);
        const T rcp = nz ? T(1) / pj : T(0);
#pragma unroll
        for (int q = 0; q < NR; q++) {
            if (act[q]) {
                T l = v[q][0];
                if (nz) l = l * rcp;
#pragma unroll
                for (int c = 1; c < TW; c++) v[q][c - 1] = fma(-l, prow[c], v[q][c]);
                v[q][TW - 1] = T(0);
                if constexpr (OUT_LU) {
#pragma unroll
                    for (int c = 1; c < TW; c++) mul[q][c - 1] = mul[q][c];
                    mul[q][TW - 1] = l;
                }
            }
        }
    }
}

// Loads candidate rows for an elimination level: NR per lane, the id of
// lane l's q-th row is ids[(q*64 + l)] (INT_MAX: none), its values come from
// A (own rows) or from the published candidates.
template <bool FULL>
__device__ __forceinline__ int colx(int c, int wa)
{
    return FULL ? c : min(c, wa - 1);
}

template <typename T, int NR, bool FROM_SLOTS, bool FULL>
__device__ __forceinline__ void load_rows(T (&v)[NR][TW], int (&key)[NR], int (&pid)[NR],
                                          const int *keys, const int *pids, int n_ids,
                                          const T *A, int lda, const T *cand_v, int wa)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const int idx = q * 64 + lane;
        const int r = idx < n_ids ? keys[idx] : INT_MAX;
        key[q] = r;
        pid[q] = (idx < n_ids && pids) ? pids[idx] : r;
        if (FROM_SLOTS) {
            const int sv = r != INT_MAX ? pid[q] : 0;
#pragma unroll
            for (int c = 0; c < TW; c++) v[q][c] = coherent_load(&cand_v[(size_t)sv * TW + c]);
        } else {
            const int rr = r != INT_MAX ? r : 0;
#pragma unroll
            for (int c = 0; c < TW; c++) v[q][c] = A[rr + (size_t)colx<FULL>(c, wa) * lda];
        }
    }
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

    // candidate lists of the levels: [wave][TW] keys (rows) and payloads
    __shared__ int s_k1[TNW * TW], s_p1[TNW * TW];
    __shared__ int s_kA[TNW / 2 * TW], s_pA[TNW / 2 * TW];
    __shared__ int s_k2[TW], s_p2[TW];
    __shared__ T s_lu[TW][TW + 1];
    __shared__ int s_x[TW];  // position of the k-th interchange's partner
    __shared__ int s_dst[2 * TW], s_src[2 * TW], s_np;

    if (!FULL)
        for (int idx = tid; idx < TW * (TW + 1); idx += TPT) s_lu[idx / (TW + 1)][idx % (TW + 1)] = T(0);
    // lists hold TW entries per wave; entries past wa stay "no row"
    for (int idx = tid; idx < TNW * TW; idx += TPT) s_k1[idx] = INT_MAX;
    if (tid < TNW / 2 * TW) s_kA[tid] = INT_MAX;
    __syncthreads();

    // 1. per-wave elimination of the wave's own rows
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
        wave_gepp<T, RPT, false, FULL>(v, key, key, wa, s_k1 + wave * TW, s_p1 + wave * TW,
                                       nullptr);
    }
    __syncthreads();

    // 2. the workgroup's TW candidates: waves 0-3 each eliminate the 2*TW rows
    //    two waves chose, then wave 0 the 4*TW survivors (original values)
    if (wave < TNW / 2) {
        T v[1][TW];
        int key[1], pid[1];
        load_rows<T, 1, false, FULL>(v, key, pid, s_k1 + wave * 2 * TW, nullptr, 2 * TW, A, lda,
                               nullptr, wa);
        wave_gepp<T, 1, false, FULL>(v, key, pid, wa, s_kA + wave * TW, s_pA + wave * TW,
                                     nullptr);
    }
    __syncthreads();
    if (wave == 0) {
        T v[2][TW];
        int key[2], pid[2];
        load_rows<T, 2, false, FULL>(v, key, pid, s_kA, nullptr, TNW / 2 * TW, A, lda, nullptr,
                                     wa);
        wave_gepp<T, 2, false, FULL>(v, key, pid, wa, s_k2, s_p2, nullptr);
    }
    __syncthreads();
    // publish the workgroup's candidates (ids, original values); workgroup 0
    // also the leaf's top rows
    for (int idx = tid; idx < TW * TW; idx += TPT) {
        const int j = idx / TW, c = idx % TW;
        const int r = (j < wa) ? s_k2[j] : INT_MAX;
        const T x = r != INT_MAX ? A[r + (size_t)colx<FULL>(c, wa) * lda] : T(0);
        coherent_store(&slots.cand_v[((size_t)g * TW + j) * TW + c], x);
        if (c == 0) coherent_store(&slots.cand_r[g * TW + j], r);
        if (g == 0) {
            const T y = (j < wa) ? A[j + (size_t)colx<FULL>(c, wa) * lda] : T(0);
            coherent_store(&slots.top_v[j * TW + c], y);
        }
    }

    // 3. device-wide barrier
    if (!grid_sync(ctr, target, abort_flag, spin_limit)) return;

    // 4. the tournament, identical in every workgroup: 8 waves on slices of
    //    the G*TW candidates, 4 waves on pairs of survivors, wave 0 on the last
    //    4*TW (its elimination of the pivot rows is L11\U11)
    for (int idx = tid; idx < TNW * TW; idx += TPT) s_k1[idx] = INT_MAX;
    if (tid < TNW / 2 * TW) s_kA[tid] = INT_MAX;
    __syncthreads();
    {
        const int total = G * TW;
        const int cpw = ceil_div(total, TNW);  // <= TMR * 64
        T v[TMR][TW];
        int key[TMR], pid[TMR];
#pragma unroll
        for (int q = 0; q < TMR; q++) {
            const int il = q * 64 + lane, sid = wave * cpw + il;
            const bool ok = il < cpw && sid < total;
            const int r = ok ? coherent_load(&slots.cand_r[sid]) : INT_MAX;
            key[q] = r;
            pid[q] = ok ? sid : -1;
            const int sv = (r != INT_MAX) ? sid : 0;
#pragma unroll
            for (int c = 0; c < TW; c++) v[q][c] = coherent_load(&slots.cand_v[(size_t)sv * TW + c]);
        }
        wave_gepp<T, TMR, false, FULL>(v, key, pid, wa, s_k1 + wave * TW, s_p1 + wave * TW,
                                       nullptr);
    }
    __syncthreads();
    if (wave < TNW / 2) {
        T v[1][TW];
        int key[1], pid[1];
        load_rows<T, 1, true, FULL>(v, key, pid, s_k1 + wave * 2 * TW, s_p1 + wave * 2 * TW, 2 * TW,
                              nullptr, lda, slots.cand_v, wa);
        wave_gepp<T, 1, false, FULL>(v, key, pid, wa, s_kA + wave * TW, s_pA + wave * TW,
                                     nullptr);
    }
    __syncthreads();
    if (wave == 0) {
        T v[2][TW];
        int key[2], pid[2];
        load_rows<T, 2, true, FULL>(v, key, pid, s_kA, s_pA, TNW / 2 * TW, nullptr, lda, slots.cand_v,
                              wa);
        wave_gepp<T, 2, true, FULL>(v, key, pid, wa, s_k2, s_p2, s_lu);
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
            if (lane == 0) {
                s_x[k] = xpos;
                if (g == 0) ipiv[row0 + k] = row0 + xpos + 1;
            }
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

    // 5c. the interchanges on the block column's other columns, in order;
    //     column e always by thread e / G of workgroup e % G
    const int ne = lcols + rcols;
    for (int e = g + G * tid; e < ne; e += G * TPT) {
        const int cc = (e < lcols) ? (e - lcols) : (wa + e - lcols);
        T *col = A + (ptrdiff_t)cc * lda;
        for (int k = 0; k < wa; k++) {
            const int xk = s_x[k];
            if (xk != k) {
                const T a = col[k], b = col[xk];
                col[k] = b;
                col[xk] = a;
            }
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
