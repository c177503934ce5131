// panel.hip -- the unblocked panel factorization PDGETF2 (SRC/pdgetf2.f:206-237)
// for a tall leaf of at most W columns, as one launch of G resident
// workgroups that hold the leaf in registers.
//
// Per column j the reference does PDAMAX (pivot search) -> PDSWAP (row
// interchange) -> PDSCAL(1/GMAX) -> PDGER(-1) (rank-1 update).  Here:
//   1. every thread scans its rows >= j; a wave64 shuffle reduction and a
//      4-wave LDS step give the workgroup's candidate (|v|, row), ties to
//      the lower row as IDAMAX does;
//   2. each workgroup publishes its candidate and that row's W values, and
//      workgroup 0 publishes row j;
//   3. one device-wide barrier;
//   4. every workgroup reduces the G candidates in the same total order, so
//      all agree on the pivot row p without a second barrier;
//   5. the owners of rows j and p exchange them in registers, and the
//      interchange is applied to the rest of the block column outside the
//      leaf (PDGETF2 swaps the full JB-wide panel rows, pdgetf2.f:219-220),
//      each of those columns always by the same thread, so successive
//      interchanges of one column stay in order;
//   6. rows below j are scaled by the reciprocal 1/pivot (the reference's
//      ONE / GMAX multiply, pdgetf2.f:224) and updated a -= l * u with one fma.
// A zero pivot skips the interchange and the scaling and records the first
// such column (pdgetf2.f:226-228); the update still runs, as PDGER does.
#include "sla_common.h"
#include "sla_kernels.h"

#include <algorithm>

#include <climits>

namespace sla {

namespace {

constexpr int PT = 512;          // threads per workgroup (8 waves)
constexpr int NWV = PT / 64;
constexpr int LEAF_GMAX = 128;   // workgroups per leaf launch at most

template <typename T, int W, int RPT>
__global__ __launch_bounds__(PT) void panel_leaf_kernel(
    T *__restrict__ A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
    int *__restrict__ ipiv, PanelSlots<T> slots, unsigned *ctr, unsigned base, int *abort_flag,
    int *zero_col, uint64_t spin_limit)
{
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int R = RPT * PT;  // rows per workgroup

    __shared__ double s_wv[NWV];
    __shared__ int s_wr[NWV];
    __shared__ double s_cv[LEAF_GMAX];
    __shared__ int s_cr[LEAF_GMAX];
    __shared__ T s_rows[LEAF_GMAX][W];
    __shared__ T s_rowj[W];
    __shared__ int s_p;

    T a[RPT][W];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = g * R + q * PT + tid;
#pragma unroll
        for (int c = 0; c < W; c++)
            a[q][c] = (r < m && c < wa) ? A[r + (size_t)c * lda] : T(0);
    }

#pragma unroll
    for (int j = 0; j < W; j++) {
        if (j < wa) {  // uniform across the grid
        const int par = j & 1;

        // 1. candidate of this workgroup: per thread, per wave (shuffles),
        //    then every thread folds the NWV wave results itself
        double bv = -1.0;
        int br = INT_MAX;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * PT + tid;
            if (r < m && r >= j) {
                const double v = fabs((double)a[q][j]);
                if (v > bv) { bv = v; br = r; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int orr = __shfl_xor(br, off);
            if (pivot_better(ov, orr, bv, br)) { bv = ov; br = orr; }
        }
        if (lane == 0) { s_wv[wave] = bv; s_wr[wave] = br; }
        __syncthreads();
        double cv = s_wv[0];
        int cr = s_wr[0];
#pragma unroll
        for (int w = 1; w < NWV; w++)
            if (pivot_better(s_wv[w], s_wr[w], cv, cr)) { cv = s_wv[w]; cr = s_wr[w]; }

        // 2. publish the candidate (value, row, the row's W entries) and,
        //    from workgroup 0, row j
        if (tid == 0) {
            coherent_store(&slots.cand_v[par * slots.gmax + g], cv);
            coherent_store(&slots.cand_r[par * slots.gmax + g], cr);
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * PT + tid;
            if (r == cr) {
                T *dst = slots.rows + ((size_t)par * slots.gmax + g) * W;
#pragma unroll
                for (int c = 0; c < W; c++) coherent_store(&dst[c], a[q][c]);
            }
            if (r == j) {
                T *dst = slots.rowj + (size_t)par * W;
#pragma unroll
                for (int c = 0; c < W; c++) coherent_store(&dst[c], a[q][c]);
            }
        }

        // 3. device-wide barrier
        if (!grid_sync(ctr, base + (unsigned)(j + 1) * (unsigned)G, abort_flag, spin_limit))
            return;

        // 4. one remote round trip: all candidates, all candidate rows, row j
        if (tid < G) {
            s_cv[tid] = coherent_load(&slots.cand_v[par * slots.gmax + tid]);
            s_cr[tid] = coherent_load(&slots.cand_r[par * slots.gmax + tid]);
        }
        for (int idx = tid; idx < G * W; idx += PT)
            s_rows[idx / W][idx % W] = coherent_load(&slots.rows[(size_t)par * slots.gmax * W + idx]);
        if (tid < W) s_rowj[tid] = coherent_load(&slots.rowj[(size_t)par * W + tid]);
        __syncthreads();

        // 5. global pivot: the same reduction order in every workgroup
        if (wave == 0) {
            double v = -1.0;
            int r = INT_MAX;
            for (int x = lane; x < G; x += 64)
                if (pivot_better(s_cv[x], s_cr[x], v, r)) { v = s_cv[x]; r = s_cr[x]; }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off);
                const int orr = __shfl_xor(r, off);
                if (pivot_better(ov, orr, v, r)) { v = ov; r = orr; }
            }
            if (lane == 0) s_p = (r == INT_MAX) ? j : r;  // an all-NaN column keeps row j
        }
        __syncthreads();
        const int p = s_p;
        const int gw = p / R;
        const T piv = s_rows[gw][j];
        const bool nonzero = (piv != T(0));

        // 6. interchange rows j and p (only if the pivot is nonzero)
        if (nonzero && p != j) {
#pragma unroll
            for (int q = 0; q < RPT; q++) {
                const int r = g * R + q * PT + tid;
                if (r == p) {
#pragma unroll
                    for (int c = 0; c < W; c++) a[q][c] = s_rowj[c];
                }
                if (r == j) {
#pragma unroll
                    for (int c = 0; c < W; c++) a[q][c] = s_rows[gw][c];
                }
            }
            // The same interchange on the block column's other columns: the
            // lcols columns left of the leaf and the rcols right of it.
            // Column e is always handled by thread e / G of workgroup e % G.
            const int ne = lcols + rcols;
            for (int e = g + G * tid; e < ne; e += G * PT) {
                const int cc = (e < lcols) ? (e - lcols) : (wa + e - lcols);
                T *col = A + (ptrdiff_t)cc * lda;
                const T x = col[j], y = col[p];
                col[j] = y;
                col[p] = x;
            }
        }
        if (g == 0 && tid == 0) {
            // PDAMAX returns INDX = IX for a zero maximum (pdamax_.c:486).
            ipiv[row0 + j] = row0 + (nonzero ? p : j) + 1;
            if (!nonzero) atomicMin(zero_col, col0 + j + 1);
        }

        // 7. scale by the reciprocal and rank-1 update of the rows below j
        const T rcp = nonzero ? T(1) / piv : T(0);
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * PT + tid;
            if (r < m && r > j) {
                T l = a[q][j];
                if (nonzero) {
                    l = l * rcp;
                    a[q][j] = l;
                }
#pragma unroll
                for (int c = j + 1; c < W; c++) a[q][c] = fma(-l, s_rows[gw][c], a[q][c]);
            }
        }
        }  // j < wa
    }

#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = g * R + q * PT + tid;
        if (r < m) {
#pragma unroll
            for (int c = 0; c < W; c++)
                if (c < wa) A[r + (size_t)c * lda] = a[q][c];
        }
    }
}

template <typename T, int W, int RPT>
void launch_leaf_rpt(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                     int *ipiv, const PanelWork<T> &pw, unsigned base, int G, hipStream_t s)
{
    // Only register-resident shapes are instantiated: RPT * W values of T in
    // at most 512 bytes per thread (128 VGPRs).
    if constexpr (RPT * W * sizeof(T) <= 512) {
        hipLaunchKernelGGL((panel_leaf_kernel<T, W, RPT>), dim3(G), dim3(PT), 0, s, A, lda, m,
                           wa, lcols, rcols, row0, col0, ipiv, pw.slots, pw.ctr, base,
                           pw.abort_flag, pw.zero_col, pw.spin_limit);
        SLA_LAUNCH_CHECK();
    } else {
        throw std::logic_error("panel leaf: rows per thread exceed the register budget");
    }
}

template <typename T, int W>
void launch_leaf_w(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                   int *ipiv, const PanelWork<T> &pw, unsigned base, int G, int rpt,
                   hipStream_t s)
{
#define SLA_RPT(R) launch_leaf_rpt<T, W, R>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, base, G, s)
    switch (rpt) {
    case 1: SLA_RPT(1); break;
    case 2: SLA_RPT(2); break;
    case 4: SLA_RPT(4); break;
    default: SLA_RPT(8); break;
    }
#undef SLA_RPT
}

}  // namespace

template <typename T>
unsigned panel_leaf(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                    int *ipiv, const PanelWork<T> &pw, unsigned base, hipStream_t s)
{
    // Rows per thread grows (1, 2, 4, 8) until the grid is within the target
    // workgroup count; the register budget caps rows/thread * W at 64
    // doubles (128 floats).
    const int max_rpt_w = (int)(512 / sizeof(T));
    const int wcap = wa <= 8 ? 8 : (wa <= 16 ? 16 : 32);
    int rpt_cap = max_rpt_w / wcap;
    if (rpt_cap > 8) rpt_cap = 8;
    // the caller's preferred grid, capped by what the leaf's exchange holds
    const int target = std::min(pw.target_wgs, std::min(LEAF_GMAX, pw.slots.gmax));
    int rpt = 1;
    int G = ceil_div(m, PT * rpt);
    while (G > target && rpt < rpt_cap) {
        rpt *= 2;
        G = ceil_div(m, PT * rpt);
    }
    if (G > pw.slots.gmax || G > LEAF_GMAX)
        throw std::runtime_error("panel leaf: panel too tall for the leaf's workgroup limit");
    if (wa <= 8)
        launch_leaf_w<T, 8>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, base, G, rpt, s);
    else if (wa <= 16)
        launch_leaf_w<T, 16>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, base, G, rpt, s);
    else if (wa <= 32)
        launch_leaf_w<T, 32>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, base, G, rpt, s);
    else
        throw std::runtime_error("panel leaf: width above 32");
    return (unsigned)wa * (unsigned)G;
}

// The build compiles this file twice, SLA_PANEL_TYPE = 64 and 32, so that
// the fully unrolled leaf instances of the two storage types build in
// parallel; without the macro both are instantiated here.
#if !defined(SLA_PANEL_TYPE) || SLA_PANEL_TYPE == 64
template unsigned panel_leaf<double>(double *, int, int, int, int, int, int, int, int *,
                                     const PanelWork<double> &, unsigned, hipStream_t);
#endif
#if !defined(SLA_PANEL_TYPE) || SLA_PANEL_TYPE == 32
template unsigned panel_leaf<float>(float *, int, int, int, int, int, int, int, int *,
                                    const PanelWork<float> &, unsigned, hipStream_t);
#endif

}  // namespace sla
