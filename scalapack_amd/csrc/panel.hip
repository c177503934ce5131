// panel.hip -- the unblocked panel factorization PDGETF2 (SRC/pdgetf2.f:206-237)
// for a tall leaf of at most W columns, as one launch of G resident
// workgroups that hold the leaf in registers.
//
// Per column j the reference does PDAMAX (pivot search) -> PDSWAP (row
// interchange) -> PDSCAL(1/GMAX) -> PDGER(-1) (rank-1 update).  Here:
//   1. every thread scans its rows >= j; a wave64 shuffle reduction and a
//      4-wave LDS step give the workgroup's candidate (|v|, row), ties to
//      the lower row as IDAMAX does;
//   2. each workgroup publishes its candidate row (index and W values) and
//      workgroup 0 publishes row j, into mailbox words of column j;
//   3. every workgroup gathers the G candidates and row j; each word is
//      written once per launch and starts out empty, so a word that is no
//      longer empty holds this column's value: one round trip, no counter;
//   4. every workgroup reduces the G candidates in the same total order, so
//      all agree on the pivot row p;
//   5. the owners of rows j and p exchange them in registers;
//   6. rows below j are scaled by the reciprocal 1/pivot (the reference's
//      ONE / GMAX multiply, pdgetf2.f:224) and updated a -= l * u with one fma.
// PDGETF2 swaps the full JB-wide panel rows (pdgetf2.f:219-220): the leaf's
// interchanges, composed into one row permutation, are applied to the rest
// of the block column after the last column (no other access touches those
// columns meanwhile, so the result is the same as swapping per column).
// A zero pivot skips the interchange and the scaling and records the first
// such column (pdgetf2.f:226-228); the update still runs, as PDGER does.
//
// Mailbox: two regions of 32 column blocks; launch L uses region L & 1 and
// refills the words launch L - 1 wrote in region (L + 1) & 1 -- that launch
// has finished -- with the empty pattern, so that launch L + 1 finds its
// region empty.
#include "sla_common.h"
#include "sla_kernels.h"

#include <algorithm>
#include <type_traits>

#include <climits>

// wavefront-wide maximum of the ROCm device library
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_max_u64(
    unsigned long long);

namespace sla {

namespace {

constexpr int PT = 512;          // threads per workgroup (8 waves)
constexpr int NWV = PT / 64;
constexpr int LEAF_GMAX = 128;   // workgroups per leaf launch at most
static_assert(LEAF_GMAX <= PANEL_MAIL_GMAX, "mailbox too small for the leaf's grid");

constexpr uint64_t MAIL_EMPTY = ~0ull;  // no row index or value word has this pattern

// Storage of one value in a mailbox word.  A double whose bits equal the
// empty pattern (a NaN) is stored as the neighbouring NaN pattern.
__device__ __forceinline__ uint64_t mail_word(double x)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    return b == MAIL_EMPTY ? MAIL_EMPTY - 1 : b;
}
__device__ __forceinline__ uint64_t mail_word(float x)
{
    return (uint64_t)__builtin_bit_cast(uint32_t, x);
}
template <typename T>
__device__ __forceinline__ T mail_value(uint64_t w)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, w);
    else return __builtin_bit_cast(T, (uint32_t)w);
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

typedef unsigned long long u64x2_t __attribute__((ext_vector_type(2)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));

// Pivot key of a candidate: the bits of |v| plus one, monotone in |v|; 0 for
// no candidate and for a NaN, which IDAMAX's comparisons never select.
__device__ __forceinline__ uint64_t mag_key(double v, bool valid)
{
    const double a = fabs(v);
    return (valid && a == a) ? __builtin_bit_cast(uint64_t, a) + 1ull : 0ull;
}

// Wave-wide first maximum: the largest key and, among the lanes holding it,
// the lowest row (INT_MAX when no lane holds a key), in every lane.
__device__ __forceinline__ void wave_argmax(uint64_t &key, int &row)
{
    const uint64_t km = __ockl_wfred_max_u64(key);
    uint64_t tied = __ballot(km != 0 && key == km);
    int best = INT_MAX;
    while (tied) {  // uniform; one lane unless magnitudes tie
        const int l = __ffsll((unsigned long long)tied) - 1;
        best = min(best, __builtin_amdgcn_readlane(row, l));
        tied &= tied - 1;
    }
    key = km;
    row = best;
}

// Gathered word idx of a column into LDS: candidate idx / (W + 1) (row
// index, then W values) for idx < ncand, else row j.
template <typename T, int W>
__device__ __forceinline__ void mail_put(int idx, int ncand, uint64_t v, int *s_cr,
                                         T (*s_rows)[W + 1], T *s_rowj)
{
    if (idx < ncand) {
        const int gg = idx / (W + 1), e = idx - gg * (W + 1);
        if (e == 0) s_cr[gg] = (int)v;
        else s_rows[gg][e - 1] = mail_value<T>(v);
    } else {
        s_rowj[idx - ncand] = mail_value<T>(v);
    }
}

// Waits until the mailbox word *p is written; false once the launch has
// been aborted (by this wait after spin_limit ticks since t0, or elsewhere).
__device__ __forceinline__ bool mail_wait(const uint64_t *p, uint64_t &v, int *abort_flag,
                                          uint64_t t0, uint64_t spin_limit)
{
    unsigned it = 0;
    while ((v = coherent_load(p)) == MAIL_EMPTY) {
        if ((++it & 15u) == 0) {
            if (coherent_load(abort_flag)) return false;
            if (wall_clock64() - t0 > spin_limit) {
                coherent_store(abort_flag, 1);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

template <typename T, int W, int RPT>
__global__ __launch_bounds__(PT) void panel_leaf_kernel(
    T *__restrict__ A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
    int *__restrict__ ipiv, uint64_t *__restrict__ mail, int region, int prev_wa, int prev_g,
    int prev_cs, int *abort_flag, int *zero_col, uint64_t spin_limit)
{
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int R = RPT * PT;  // rows per workgroup
    constexpr int CS = W + 1;    // words per candidate: row index, W values
    // gathered words per thread: G * CS candidate words + W words of row j
    constexpr int KW = (LEAF_GMAX * CS + W + PT - 1) / PT;

    __shared__ __align__(16) uint64_t s_wk[NWV];
    __shared__ __align__(16) int s_wr[NWV];
    __shared__ int s_cr[LEAF_GMAX];
    __shared__ T s_rows[LEAF_GMAX][W + 1];  // odd stride: column j of all rows without bank conflicts
    __shared__ T s_rowj[W];
    __shared__ int s_abort;
    __shared__ int s_pv[W];              // interchange of each column (p; j when none)
    __shared__ int s_dst[2 * W], s_src[2 * W], s_npair;

    if (tid == 0) s_abort = 0;

    // empty the words the previous launch wrote (the other region):
    // columns < prev_wa, its prev_g candidates of prev_cs words, row j
    {
        uint64_t *o = mail + (size_t)(region ^ 1) * PANEL_MAIL_REGION;
        const int per = prev_g * prev_cs + 32;
        for (int i = g * PT + tid; i < prev_wa * per; i += G * PT) {
            const int jj = i / per, e = i - jj * per;
            const int off = e < prev_g * prev_cs ? e : PANEL_MAIL_ROWJ + (e - prev_g * prev_cs);
            coherent_store(&o[(size_t)jj * PANEL_MAIL_COL + off], MAIL_EMPTY);
        }
    }
    uint64_t *box = mail + (size_t)region * PANEL_MAIL_REGION;

    T a[RPT][W];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = g * R + q * PT + tid;
#pragma unroll
        for (int c = 0; c < W; c++)
            a[q][c] = (r < m && c < wa) ? A[r + (size_t)c * lda] : T(0);
    }
    const uint64_t t0 = wall_clock64();

    // The columns as a compile-time loop: every index into a[][] is a
    // constant, so the leaf stays in registers.  (A run-time column loop
    // with selects instead of constant indices measured slower.)
    bool alive = true;
    static_for<0, W>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (!alive || j >= wa) return;  // wa and the abort flag are uniform across the grid
        uint64_t *cb = box + (size_t)j * PANEL_MAIL_COL;  // this column's words

        // 1. candidate of this workgroup: per thread, per wave (packed
        //    magnitude keys, one wave reduction), then every thread folds
        //    the NWV wave results itself
        uint64_t bk = 0;
        int br = INT_MAX;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * PT + tid;
            const uint64_t k = mag_key((double)a[q][j], r < m && r >= j);
            if (k > bk) { bk = k; br = r; }
        }
        wave_argmax(bk, br);
        if (lane == 0) { s_wk[wave] = bk; s_wr[wave] = br; }
        __syncthreads();
        // all NWV posts in one round of 16-byte LDS reads, then a
        // branch-free fold in registers
        uint64_t wk[NWV];
        int wr[NWV];
#pragma unroll
        for (int w = 0; w < NWV; w += 2) {
            const u64x2_t k2 = *reinterpret_cast<const u64x2_t *>(&s_wk[w]);
            wk[w] = k2[0];
            wk[w + 1] = k2[1];
        }
#pragma unroll
        for (int w = 0; w < NWV; w += 4) {
            const i32x4_t r4 = *reinterpret_cast<const i32x4_t *>(&s_wr[w]);
#pragma unroll
            for (int e = 0; e < 4; e++) wr[w + e] = r4[e];
        }
        uint64_t ck = wk[0];
        int cr = wr[0];
#pragma unroll
        for (int w = 1; w < NWV; w++) {
            const bool take = wk[w] > ck || (wk[w] == ck && wr[w] < cr);
            ck = take ? wk[w] : ck;
            cr = take ? wr[w] : cr;
        }

        // 2. publish the candidate (row index and the row's W values; a
        //    workgroup without rows >= j publishes INT_MAX and zeros) and,
        //    from the owner of row j, row j
        uint64_t *mine = cb + (size_t)g * CS;
        if (cr == INT_MAX && tid == 0) {
#pragma unroll
            for (int c = 0; c < W; c++) coherent_store(&mine[1 + c], mail_word(T(0)));
            coherent_store(&mine[0], (uint64_t)INT_MAX);
        }
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = g * R + q * PT + tid;
            if (r == cr) {
#pragma unroll
                for (int c = 0; c < W; c++) coherent_store(&mine[1 + c], mail_word(a[q][c]));
                coherent_store(&mine[0], (uint64_t)cr);
            }
            if (r == j) {
                uint64_t *dst = cb + PANEL_MAIL_ROWJ;
#pragma unroll
                for (int c = 0; c < W; c++) coherent_store(&dst[c], mail_word(a[q][c]));
            }
        }

        // 3. gather: all loads in flight at once; a word still empty is
        //    then awaited on its own (bounded wait)
        {
            const int nword = G * CS + W;
            unsigned miss = 0;  // bit k: word tid + k PT not written yet
#pragma unroll
            for (int k = 0; k < KW; k++) {
                const int idx = tid + k * PT;
                if (idx < nword) {
                    const int off = idx < G * CS ? idx : PANEL_MAIL_ROWJ + (idx - G * CS);
                    const uint64_t v = coherent_load(&cb[off]);
                    if (v == MAIL_EMPTY) miss |= 1u << k;
                    else mail_put<T, W>(idx, G * CS, v, s_cr, s_rows, s_rowj);
                }
            }
            while (miss) {
                const int idx = tid + (__ffs(miss) - 1) * PT;
                const int off = idx < G * CS ? idx : PANEL_MAIL_ROWJ + (idx - G * CS);
                uint64_t v;
                if (!mail_wait(&cb[off], v, abort_flag, t0, spin_limit)) {
                    s_abort = 1;
                    break;
                }
                mail_put<T, W>(idx, G * CS, v, s_cr, s_rows, s_rowj);
                miss &= miss - 1;
            }
        }
        __syncthreads();
        if (s_abort) {
            alive = false;
            return;
        }

        // 4. global pivot: every wave reduces the G candidates itself (the
        //    same result everywhere, no further barrier); a candidate's value
        //    is |its row's entry in column j|
        int p;
        {
            uint64_t k = 0;
            int r = INT_MAX;
            for (int x = lane; x < G; x += 64) {
                const int xr = s_cr[x];
                const uint64_t xk = mag_key((double)s_rows[x][j], xr != INT_MAX);
                if (xk > k || (xk == k && xr < r)) { k = xk; r = xr; }
            }
            wave_argmax(k, r);
            p = (r == INT_MAX) ? j : r;  // an all-NaN column keeps row j
        }
        const int gw = p / R;
        const T piv = s_rows[gw][j];
        const bool nonzero = (piv != T(0));

        // 5. interchange rows j and p in registers (only if the pivot is
        //    nonzero); the rest of the block column follows after the leaf
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
        }
        if (tid == 0) s_pv[j] = nonzero ? p : j;
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
                for (int c = j + 1; c < W; c++) a[q][c] = fma(-l, s_rows[gw][c], a[q][c]);
            }
        }
    });
    if (!alive) return;

#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = g * R + q * PT + tid;
        if (r < m) {
#pragma unroll
            for (int c = 0; c < W; c++)
                if (c < wa) A[r + (size_t)c * lda] = a[q][c];
        }
    }

    // The leaf's interchanges on the block column's other columns: the
    // lcols columns left of the leaf and the rcols right of it.  Wave 0
    // composes the sequence (j, p_j) into destination <- source row pairs
    // (lane k holds pair k; at most 2 wa <= 64 rows move).  Then every wave
    // rewrites CR columns at a time, lane k moving pair k: the wave's loads
    // of a column all complete before its stores to that column.
    const int ne = lcols + rcols;
    if (ne <= 0) return;
    __syncthreads();  // s_pv complete
    if (wave == 0) {
        int dst = -1, src = -1, cnt = 0;
        for (int j = 0; j < wa; j++) {
            const int p = s_pv[j];
            if (p == j) continue;
            const uint64_t mj = __ballot(dst == j), mp = __ballot(dst == p);
            const int lj = mj ? __ffsll((unsigned long long)mj) - 1 : -1;
            const int lp = mp ? __ffsll((unsigned long long)mp) - 1 : -1;
            const int sj = lj >= 0 ? __shfl(src, lj) : j;  // row now at position j
            const int sp = lp >= 0 ? __shfl(src, lp) : p;  // row now at position p
            if (lj >= 0) { if (lane == lj) src = sp; }
            else { if (lane == cnt) { dst = j; src = sp; } cnt++; }
            if (lp >= 0) { if (lane == lp) src = sj; }
            else { if (lane == cnt) { dst = p; src = sj; } cnt++; }
        }
        if (lane < cnt) { s_dst[lane] = dst; s_src[lane] = src; }
        if (lane == 0) s_npair = cnt;
    }
    __syncthreads();
    if (lane >= s_npair) return;
    constexpr int CR = 8;
    const int rd = s_dst[lane], rs = s_src[lane];
    for (int e0 = (g * NWV + wave) * CR; e0 < ne; e0 += G * NWV * CR) {
        T v[CR];
#pragma unroll
        for (int c = 0; c < CR; c++) {
            const int e = e0 + c;
            const int cc = (e < lcols) ? (e - lcols) : (wa + e - lcols);
            if (e < ne) v[c] = A[(ptrdiff_t)cc * lda + rs];
        }
#pragma unroll
        for (int c = 0; c < CR; c++) {
            const int e = e0 + c;
            const int cc = (e < lcols) ? (e - lcols) : (wa + e - lcols);
            if (e < ne) A[(ptrdiff_t)cc * lda + rd] = v[c];
        }
    }
}

template <typename T, int W, int RPT>
void launch_leaf_rpt(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                     int *ipiv, const PanelWork<T> &pw, int G, hipStream_t s)
{
    // Only register-resident shapes are instantiated: at most 64 values of
    // T per thread (128 VGPRs for doubles; floats get the same cap, beyond
    // which the leaf's other registers no longer fit without scratch).
    if constexpr (RPT * W <= 64) {
        PanelMailState &ms = *pw.mail_state;
        hipLaunchKernelGGL((panel_leaf_kernel<T, W, RPT>), dim3(G), dim3(PT), 0, s, A, lda, m,
                           wa, lcols, rcols, row0, col0, ipiv, pw.mail, (int)(ms.launches & 1u),
                           ms.prev_wa, ms.prev_g, ms.prev_cs, pw.abort_flag, pw.zero_col,
                           pw.spin_limit);
        SLA_LAUNCH_CHECK();
        ms.launches++;
        ms.prev_wa = wa;
        ms.prev_g = G;
        ms.prev_cs = W + 1;
    } else {
        throw std::logic_error("panel leaf: rows per thread exceed the register budget");
    }
}

template <typename T, int W>
void launch_leaf_w(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                   int *ipiv, const PanelWork<T> &pw, int G, int rpt,
                   hipStream_t s)
{
#define SLA_RPT(R) launch_leaf_rpt<T, W, R>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, G, s)
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
                    int *ipiv, const PanelWork<T> &pw, hipStream_t s)
{
    // Rows per thread grows (1, 2, 4, 8) until the grid is within the target
    // workgroup count; the register budget caps rows/thread * W at 64.
    const int max_rpt_w = 64;
    const int wcap = wa <= 8 ? 8 : (wa <= 16 ? 16 : 32);
    int rpt_cap = max_rpt_w / wcap;
    if (rpt_cap > 8) rpt_cap = 8;
    // the caller's preferred grid, capped by what the leaf's exchange holds
    const int target = std::min(pw.target_wgs, LEAF_GMAX);
    int rpt = 1;
    int G = ceil_div(m, PT * rpt);
    while (G > target && rpt < rpt_cap) {
        rpt *= 2;
        G = ceil_div(m, PT * rpt);
    }
    if (G > LEAF_GMAX)
        throw std::runtime_error("panel leaf: panel too tall for the leaf's workgroup limit");
    if (wa <= 8)
        launch_leaf_w<T, 8>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, G, rpt, s);
    else if (wa <= 16)
        launch_leaf_w<T, 16>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, G, rpt, s);
    else if (wa <= 32)
        launch_leaf_w<T, 32>(A, lda, m, wa, lcols, rcols, row0, col0, ipiv, pw, G, rpt, s);
    else
        throw std::runtime_error("panel leaf: width above 32");
    return 0;  // the mailbox exchange adds no barrier arrivals
}

// The build compiles this file twice, SLA_PANEL_TYPE = 64 and 32, so that
// the fully unrolled leaf instances of the two storage types build in
// parallel; without the macro both are instantiated here.
#if !defined(SLA_PANEL_TYPE) || SLA_PANEL_TYPE == 64
template unsigned panel_leaf<double>(double *, int, int, int, int, int, int, int, int *,
                                     const PanelWork<double> &, hipStream_t);
#endif
#if !defined(SLA_PANEL_TYPE) || SLA_PANEL_TYPE == 32
template unsigned panel_leaf<float>(float *, int, int, int, int, int, int, int, int *,
                                    const PanelWork<float> &, hipStream_t);
#endif

}  // namespace sla
