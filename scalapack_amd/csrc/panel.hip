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
//   5. the owners of rows j and p exchange them in registers;
//   6. rows below j are scaled by the reciprocal 1/pivot (the reference's
//      ONE / GMAX multiply, pdgetf2.f:224) and updated a -= l * u with one fma.
// A zero pivot skips the interchange and the scaling and records the first
// such column (pdgetf2.f:226-228); the update still runs, as PDGER does.
#include "sla_common.h"
#include "sla_kernels.h"

#include <climits>

namespace sla {

namespace {

constexpr int PT = 256;  // threads per workgroup

template <typename T, int W, int RPT>
__global__ __launch_bounds__(PT) void panel_leaf_kernel(
    T *__restrict__ A, int lda, int m, int wa, int row0, int col0, int *__restrict__ ipiv,
    PanelSlots<T> slots, unsigned *ctr, unsigned base, int *abort_flag, int *zero_col,
    uint64_t spin_limit)
{
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int R = RPT * PT;  // rows per workgroup

    __shared__ double s_wv[PT / 64];
    __shared__ int s_wr[PT / 64];
    __shared__ int s_p;
    __shared__ T s_piv[W];
    __shared__ T s_rowj[W];

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

        // 1. candidate of this workgroup
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
        if (tid == 0) {
            double v = s_wv[0];
            int r = s_wr[0];
#pragma unroll
            for (int w = 1; w < PT / 64; w++)
                if (pivot_better(s_wv[w], s_wr[w], v, r)) { v = s_wv[w]; r = s_wr[w]; }
            coherent_store(&slots.cand_v[par * slots.gmax + g], v);
            coherent_store(&slots.cand_r[par * slots.gmax + g], r);
            s_p = r;
        }
        __syncthreads();

        // 2. publish the candidate row and (workgroup 0) row j
        {
            const int cr = s_p;
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
        }

        // 3. device-wide barrier
        if (!grid_sync(ctr, base + (unsigned)(j + 1) * (unsigned)G, abort_flag, spin_limit))
            return;

        // 4. global pivot: identical reduction order in every workgroup
        if (wave == 0) {
            double v = -1.0;
            int r = INT_MAX;
            for (int x = lane; x < G; x += 64) {
                const double cv = coherent_load(&slots.cand_v[par * slots.gmax + x]);
                const int cr = coherent_load(&slots.cand_r[par * slots.gmax + x]);
                if (pivot_better(cv, cr, v, r)) { v = cv; r = cr; }
            }
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
        if (tid < W) {
            const int gw = p / R;
            s_piv[tid] = coherent_load(&slots.rows[((size_t)par * slots.gmax + gw) * W + tid]);
            s_rowj[tid] = coherent_load(&slots.rowj[(size_t)par * W + tid]);
        }
        __syncthreads();
        const T piv = s_piv[j];
        const bool nonzero = (piv != T(0));

        // 5. interchange rows j and p (only if the pivot is nonzero)
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
                    for (int c = 0; c < W; c++) a[q][c] = s_piv[c];
                }
            }
        }
        if (g == 0 && tid == 0) {
            // PDAMAX returns INDX = IX for a zero maximum (pdamax_.c:486).
            ipiv[row0 + j] = row0 + (nonzero ? p : j) + 1;
            if (!nonzero) atomicMin(zero_col, col0 + j + 1);
        }

        // 6. scale by the reciprocal and rank-1 update of the rows below j
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
                for (int c = j + 1; c < W; c++) a[q][c] = fma(-l, s_piv[c], a[q][c]);
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
void launch_leaf_rpt(T *A, int lda, int m, int wa, int row0, int col0, int *ipiv,
                     const PanelWork<T> &pw, unsigned base, int G, hipStream_t s)
{
    // Only register-resident shapes are instantiated: RPT * W values of T in
    // at most 512 bytes per thread (128 VGPRs).
    if constexpr (RPT * W * sizeof(T) <= 512) {
        hipLaunchKernelGGL((panel_leaf_kernel<T, W, RPT>), dim3(G), dim3(PT), 0, s, A, lda, m,
                           wa, row0, col0, ipiv, pw.slots, pw.ctr, base, pw.abort_flag,
                           pw.zero_col, pw.spin_limit);
        SLA_LAUNCH_CHECK();
    } else {
        throw std::logic_error("panel leaf: rows per thread exceed the register budget");
    }
}

template <typename T, int W>
void launch_leaf_w(T *A, int lda, int m, int wa, int row0, int col0, int *ipiv,
                   const PanelWork<T> &pw, unsigned base, int G, int rpt, hipStream_t s)
{
    switch (rpt) {
    case 1: launch_leaf_rpt<T, W, 1>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, s); break;
    case 2: launch_leaf_rpt<T, W, 2>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, s); break;
    case 4: launch_leaf_rpt<T, W, 4>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, s); break;
    default: launch_leaf_rpt<T, W, 8>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, s); break;
    }
}

}  // namespace

template <typename T>
unsigned panel_leaf(T *A, int lda, int m, int wa, int row0, int col0, int *ipiv,
                    const PanelWork<T> &pw, unsigned base, hipStream_t s)
{
    // Rows per thread grows (1, 2, 4, 8) until the grid is within the target
    // workgroup count; the register budget caps rows/thread * W at 64
    // doubles (128 floats).
    const int max_rpt_w = (int)(512 / sizeof(T));
    const int wcap = wa <= 8 ? 8 : (wa <= 16 ? 16 : 32);
    int rpt_cap = max_rpt_w / wcap;
    if (rpt_cap > 8) rpt_cap = 8;
    int rpt = 1;
    int G = ceil_div(m, PT * rpt);
    while (G > pw.target_wgs && rpt < rpt_cap) {
        rpt *= 2;
        G = ceil_div(m, PT * rpt);
    }
    if (G > pw.slots.gmax)
        throw std::runtime_error("panel leaf: panel too tall for the configured workgroup slots");
    if (wa <= 8)
        launch_leaf_w<T, 8>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, rpt, s);
    else if (wa <= 16)
        launch_leaf_w<T, 16>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, rpt, s);
    else if (wa <= 32)
        launch_leaf_w<T, 32>(A, lda, m, wa, row0, col0, ipiv, pw, base, G, rpt, s);
    else
        throw std::runtime_error("panel leaf: width above 32");
    return (unsigned)wa * (unsigned)G;
}

template unsigned panel_leaf<double>(double *, int, int, int, int, int, int *,
                                     const PanelWork<double> &, unsigned, hipStream_t);
template unsigned panel_leaf<float>(float *, int, int, int, int, int, int *,
                                    const PanelWork<float> &, unsigned, hipStream_t);

}  // namespace sla
