// lu.hip -- host orchestration of the device LU factorization and solve.
//
// dgetrf follows the reference's right-looking blocked algorithm
// (SRC/pdgetrf.f:227-295): for each block column J of width NB
//   PDGETF2 on A(J:M, J:J+JB)                  -> panel()  (recursive, below)
//   PDLASWP on the columns left and right of it -> laswp_pairs + laswp_apply
//   PDTRSM  U12 <- L11^{-1} A12                  -> trsm_llnu
//   PDGEMM  A22 <- A22 - L21 U12                 -> dgemm_nn_sub (MFMA)
// The panel is factored recursively: halves of the column block are split
// until leaves of <= 32 columns, which one device-wide-barrier kernel
// factors; between halves the same swap / trsm / gemm steps update the right
// half.  Only the rounding order differs from the reference's column-by-column
// PDGETF2; the pivot rule (largest |a|, lowest row on ties) is the same.
//
// dgetrs follows SRC/pdgetrs.f:255-266: the row interchanges of B (PDLAPIV)
// as one gather by the factorization's permutation, then the unit-lower and
// the upper solve, each one persistent launch.
#include "sla_common.h"
#include "sla_kernels.h"
#include "sla_handle.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace sla {

namespace {

template <typename T>
T *alloc_dev(size_t n)
{
    void *p = nullptr;
    SLA_HIP(hipMalloc(&p, n * sizeof(T) + 256));
    return static_cast<T *>(p);
}

}  // namespace

// ---------------------------------------------------------------------------
Handle::Handle()
{
    SLA_HIP(hipGetDevice(&device));
    SLA_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    stream = own_stream;

    int clk_khz = 100000;
    if (hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        clk_khz <= 0)
        clk_khz = 100000;
    // A device-wide barrier or dependency wait gives up after ~4 s.
    spin_limit = (uint64_t)clk_khz * 1000ull * 4ull;

    ctl = alloc_dev<int>(64);
    ctr = reinterpret_cast<unsigned *>(ctl + 0);
    abort_flag = ctl + 1;
    zero_col = ctl + 2;
    ticket = ctl + 3;
    pair_count = ctl + 4;
    bad_pivot = ctl + 5;

    constexpr int GMAX = 256;
    slot_cand_v = alloc_dev<double>(2 * GMAX);
    slot_cand_r = alloc_dev<int>(2 * GMAX);
    slot_rows = alloc_dev<double>(2 * GMAX * 32);
    slot_rowj = alloc_dev<double>(2 * 32);
    slot_gmax = GMAX;

    pairs = alloc_dev<int>(4 * 512);

    trsv_max_blocks = 1 << 16;
    trsv_flags = alloc_dev<int>(trsv_max_blocks);
    fill_int(trsv_flags, trsv_max_blocks, 0, stream);
    fill_int(ctl, 64, 0, stream);
    SLA_HIP(hipStreamSynchronize(stream));
}

Handle::~Handle()
{
    hipStreamSynchronize(stream);
    for (void *p : {(void *)ctl, (void *)slot_cand_v, (void *)slot_cand_r, (void *)slot_rows,
                    (void *)slot_rowj, (void *)pairs, (void *)trsv_flags, (void *)perm,
                    (void *)work})
        if (p) hipFree(p);
    for (auto &e : ev_pool) hipEventDestroy(e);
    hipStreamDestroy(own_stream);
}

template <typename T>
PanelWork<T> Handle::panel_work()
{
    PanelWork<T> pw;
    pw.slots.cand_v = slot_cand_v;
    pw.slots.cand_r = slot_cand_r;
    pw.slots.rows = reinterpret_cast<T *>(slot_rows);
    pw.slots.rowj = reinterpret_cast<T *>(slot_rowj);
    pw.slots.gmax = slot_gmax;
    pw.ctr = ctr;
    pw.abort_flag = abort_flag;
    pw.zero_col = zero_col;
    pw.spin_limit = spin_limit;
    pw.target_wgs = panel_wgs;
    return pw;
}

TrsvWork Handle::trsv_work()
{
    TrsvWork tw;
    tw.ticket = ticket;
    tw.flags = trsv_flags;
    tw.epoch = trsv_epoch;
    tw.max_blocks = trsv_max_blocks;
    tw.abort_flag = abort_flag;
    tw.spin_limit = spin_limit;
    return tw;
}

void Handle::ensure_perm(int m)
{
    if (m <= perm_cap) return;
    if (perm) SLA_HIP(hipFree(perm));
    perm = alloc_dev<int>((size_t)m);
    perm_cap = m;
}

void Handle::ensure_work(size_t bytes)
{
    if (bytes <= work_cap) return;
    SLA_HIP(hipStreamSynchronize(stream));
    if (work) SLA_HIP(hipFree(work));
    work = alloc_dev<char>(bytes);
    work_cap = bytes;
}

// --- per-kernel timing -----------------------------------------------------
void Handle::timing_begin(int cat)
{
    if (!profile) return;
    hipEvent_t e0, e1;
    if (ev_free.size() < 2) {
        SLA_HIP(hipEventCreate(&e0));
        SLA_HIP(hipEventCreate(&e1));
        ev_pool.push_back(e0);
        ev_pool.push_back(e1);
    } else {
        e0 = ev_free.back();
        ev_free.pop_back();
        e1 = ev_free.back();
        ev_free.pop_back();
    }
    SLA_HIP(hipEventRecord(e0, stream));
    pending.push_back({cat, 0.0, e0, e1});
}

void Handle::timing_end(double work_units)
{
    if (!profile) return;
    Pending &p = pending.back();
    p.work = work_units;
    SLA_HIP(hipEventRecord(p.e1, stream));
}

void Handle::timing_collect()
{
    if (!profile) return;
    SLA_HIP(hipStreamSynchronize(stream));
    for (auto &p : pending) {
        float ms = 0.f;
        SLA_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
        stats[p.cat].ms += ms;
        stats[p.cat].work += p.work;
        stats[p.cat].launches += 1;
        ev_free.push_back(p.e0);
        ev_free.push_back(p.e1);
    }
    pending.clear();
}

void Handle::reset_stats()
{
    for (auto &s : stats) s = Stat{};
}

// ---------------------------------------------------------------------------
namespace {

template <typename T>
struct Lu {
    Handle &h;
    T *A;
    int lda, M, N;
    int *ipiv;  // device, indexed by global row (0-based), 1-based values
    PanelWork<T> pw;

    T *at(int i, int j) const { return A + (size_t)i + (size_t)j * lda; }

    void swap_block(int k1, int k2, T *cols, int ncols)
    {
        if (ncols <= 0) return;
        h.timing_begin(CAT_LASWP);
        laswp_apply<T>(cols, lda, 0, ncols, h.pairs, h.pair_count, 2 * (k2 - k1), h.stream);
        h.timing_end(32.0 * (double)(k2 - k1) * ncols);
    }

    void make_pairs(int k1, int k2)
    {
        h.timing_begin(CAT_LASWP);
        laswp_pairs(ipiv, k1, k2, M, h.pairs, h.pair_count, h.bad_pivot, h.stream);
        h.timing_end(0.0);
    }

    void gemm(int m, int n, int k, const T *a, const T *b, T *c);
    void trsm(int m, int n, const T *l, T *b)
    {
        if (m <= 0 || n <= 0) return;
        h.timing_begin(CAT_TRSM);
        // trsm_llnu holds at most 256 rows of L; taller blocks are split
        // into 256-row diagonal solves and a gemm update between them.
        for (int k = 0; k < m; k += 256) {
            const int kb = std::min(256, m - k);
            trsm_llnu<T>(kb, n, l + (size_t)k + (size_t)k * lda, lda, b + k, lda, h.stream);
            if (k + kb < m) gemm(m - k - kb, n, kb, l + (size_t)(k + kb) + (size_t)k * lda, b + k, b + k + kb);
        }
        h.timing_end((double)m * m * n);
    }

    // Recursive panel on rows r0.., columns c0 .. c0+n-1 (global indices).
    void panel(int r0, int c0, int n)
    {
        const int m = M - r0;
        if (n <= h.leaf_w) {
            h.timing_begin(CAT_PANEL);
            h.ctr_base += panel_leaf<T>(at(r0, c0), lda, m, n, r0, c0, ipiv, pw, h.ctr_base, h.stream);
            h.timing_end((double)m * n * n);
            return;
        }
        int n1 = (n / 2 / h.leaf_w) * h.leaf_w;
        if (n1 < h.leaf_w) n1 = h.leaf_w;
        panel(r0, c0, n1);
        make_pairs(r0, r0 + n1);
        swap_block(r0, r0 + n1, at(0, c0 + n1), n - n1);
        trsm(n1, n - n1, at(r0, c0), at(r0, c0 + n1));
        gemm(m - n1, n - n1, n1, at(r0 + n1, c0), at(r0, c0 + n1), at(r0 + n1, c0 + n1));
        panel(r0 + n1, c0 + n1, n - n1);
        make_pairs(r0 + n1, r0 + n);
        swap_block(r0 + n1, r0 + n, at(0, c0), n1);
    }

    int run()
    {
        const int mn = std::min(M, N);
        for (int j = 0; j < mn; j += h.nb) {
            const int jb = std::min(h.nb, mn - j);
            panel(j, j, jb);
            // The panel's pivots on the columns left and right of it and on
            // the row permutation kept for the solve (PDLASWP, pdgetrf.f:268,275).
            for (int k = j; k < j + jb; k += 512) {
                const int k2 = std::min(j + jb, k + 512);
                make_pairs(k, k2);
                swap_block(k, k2, at(0, 0), j);
                swap_block(k, k2, at(0, j + jb), N - j - jb);
                laswp_apply<int>(h.perm, M, 0, 1, h.pairs, h.pair_count, 2 * (k2 - k), h.stream);
            }
            if (j + jb < N) {
                trsm(jb, N - j - jb, at(j, j), at(j, j + jb));
                if (j + jb < M)
                    gemm(M - j - jb, N - j - jb, jb, at(j + jb, j), at(j, j + jb), at(j + jb, j + jb));
            }
        }
        return 0;
    }
};

template <>
void Lu<double>::gemm(int m, int n, int k, const double *a, const double *b, double *c)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    h.timing_begin(CAT_GEMM);
    dgemm_nn_sub(m, n, k, a, lda, b, lda, c, lda, h.stream);
    h.timing_end(2.0 * m * n * k);
}

}  // namespace

int dgetrf_device(Handle &h, int m, int n, double *dA, int lda, int *dipiv)
{
    if (m <= 0 || n <= 0) return 0;
    const int mn = std::min(m, n);
    h.ensure_perm(m);
    h.ctr_base = 0;
    fill_int(h.ctl, 6, 0, h.stream);         // counter, flags
    fill_int(h.zero_col, 1, INT_MAX, h.stream);
    iota_int(h.perm, m, 0, h.stream);
    Lu<double> lu{h, dA, lda, m, n, dipiv, h.panel_work<double>()};
    lu.run();
    int flags[3];
    SLA_HIP(hipMemcpyAsync(flags, h.ctl, sizeof flags, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    h.timing_collect();
    if (flags[1]) throw std::runtime_error("dgetrf: a device-wide barrier timed out (panel)");
    h.perm_key_a = dA;
    h.perm_key_ipiv = dipiv;
    h.perm_m = m;
    const int zc = flags[2];
    return (zc != INT_MAX && zc <= mn) ? zc : 0;
}

// Row permutation of the factorization from IPIV alone (for a solve that did
// not follow this handle's own factorization).
static void perm_from_ipiv(Handle &h, int n, const int *dipiv)
{
    h.ensure_perm(n);
    iota_int(h.perm, n, 0, h.stream);
    for (int k = 0; k < n; k += 512) {
        const int k2 = std::min(n, k + 512);
        laswp_pairs(dipiv, k, k2, n, h.pairs, h.pair_count, h.bad_pivot, h.stream);
        laswp_apply<int>(h.perm, n, 0, 1, h.pairs, h.pair_count, 2 * (k2 - k), h.stream);
    }
}

int dgetrs_device(Handle &h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
                  double *dB, int ldb)
{
    if (n <= 0 || nrhs <= 0) return 0;
    fill_int(h.bad_pivot, 1, 0, h.stream);
    if (!(h.perm_key_a == dA && h.perm_key_ipiv == dipiv && h.perm_m == n))
        perm_from_ipiv(h, n, dipiv);
    const int chunk = std::min(nrhs, 16);
    h.ensure_work((size_t)n * chunk * sizeof(double));
    double *W = reinterpret_cast<double *>(h.work);
    TrsvWork tw = h.trsv_work();
    fill_int(h.abort_flag, 1, 0, h.stream);
    for (int j = 0; j < nrhs; j += 16) {
        const int nr = std::min(16, nrhs - j);
        double *Bj = dB + (size_t)j * ldb;
        h.timing_begin(CAT_LASWP);
        gather_rows<double>(Bj, ldb, W, n, h.perm, n, nr, h.stream);
        h.timing_end(16.0 * n * nr);
        h.timing_begin(CAT_TRSV);
        trsv_persistent<double>(true, n, nr, dA, lda, W, n, tw, h.stream);
        h.timing_end(4.0 * (double)n * n);
        h.timing_begin(CAT_TRSV);
        trsv_persistent<double>(false, n, nr, dA, lda, W, n, tw, h.stream);
        h.timing_end(4.0 * (double)n * n);
        SLA_HIP(hipMemcpy2DAsync(Bj, (size_t)ldb * sizeof(double), W, (size_t)n * sizeof(double),
                                 (size_t)n * sizeof(double), nr, hipMemcpyDeviceToDevice, h.stream));
    }
    h.trsv_epoch = tw.epoch;
    int flags[6];
    SLA_HIP(hipMemcpyAsync(flags, h.ctl, sizeof flags, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    h.timing_collect();
    if (flags[5]) throw std::runtime_error("dgetrs: IPIV holds a row outside 1..N");
    if (flags[1]) throw std::runtime_error("dgetrs: a dependency wait timed out (triangular solve)");
    return 0;
}

void laswp_rows_device(Handle &h, double *dA, int lda, int ncols, int k1, int k2,
                       const int *dipiv, bool forward)
{
    std::vector<std::pair<int, int>> chunks;
    for (int k = k1; k < k2; k += 512) chunks.push_back({k, std::min(k2, k + 512)});
    if (!forward) std::reverse(chunks.begin(), chunks.end());
    fill_int(h.bad_pivot, 1, 0, h.stream);
    for (auto &c : chunks) {
        laswp_pairs(dipiv, c.first, c.second, lda, h.pairs, h.pair_count, h.bad_pivot,
                    h.stream);
        laswp_apply<double>(dA, lda, 0, ncols, h.pairs, h.pair_count, 2 * (c.second - c.first),
                            h.stream, !forward);
    }
    int bad = 0;
    SLA_HIP(hipMemcpyAsync(&bad, h.bad_pivot, sizeof bad, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    if (bad) throw std::runtime_error("pdlaswp: IPIV holds a row outside the matrix");
}

void laswp_cols_device(Handle &h, double *dA, int lda, int nrows, int k1, int k2,
                       const int *dipiv, bool forward)
{
    laswp_cols(dA, lda, nrows, k1, k2, dipiv, forward, h.stream);
    SLA_HIP(hipStreamSynchronize(h.stream));
}

}  // namespace sla
