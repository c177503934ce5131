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
#include <string>
#include <type_traits>
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
    if (hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess || num_cus <= 0)
        num_cus = 256;
    SLA_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    SLA_HIP(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
    stream = own_stream;
    SLA_HIP(hipEventCreateWithFlags(&ev_next, hipEventDisableTiming));
    SLA_HIP(hipEventCreateWithFlags(&ev_panel, hipEventDisableTiming));
    SLA_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));

    int clk_khz = 100000;
    if (hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        clk_khz <= 0)
        clk_khz = 100000;
    wall_khz = clk_khz;
    // A device-wide barrier or dependency wait gives up after ~4 s.
    spin_limit = (uint64_t)clk_khz * 1000ull * 4ull;

    ctl = alloc_dev<int>(64);
    ctr = reinterpret_cast<unsigned *>(ctl + 0);
    abort_flag = ctl + 1;
    zero_col = ctl + 2;
    ticket = ctl + 3;
    pair_count = ctl + 4;
    bad_pivot = ctl + 5;

    panel_mail = alloc_dev<uint64_t>(PANEL_MAIL_WORDS);
    fill_u32(reinterpret_cast<unsigned *>(panel_mail), (int)(2 * PANEL_MAIL_WORDS), ~0u, stream);

    pairs = alloc_dev<int>(2 * 2048);
    pair_counts = alloc_dev<int>(2);

    fill_int(ctl, 64, 0, stream);
    kstamps = alloc_dev<unsigned long long>(2 * KSTAMP_CAP);
    reset_kstamps(kstamps, KSTAMP_CAP, stream);
    SLA_HIP(hipStreamSynchronize(stream));
}

Handle::~Handle()
{
    (void)hipStreamSynchronize(stream);
    batched_plan_free(batched);
    for (void *p : {(void *)ctl, (void *)panel_mail, (void *)pairs, (void *)pair_counts,
                    (void *)perm, (void *)work, (void *)kstamps})
        if (p) hipFree(p);
    for (auto &b : bufs)
        if (b.p) hipFree(b.p);
    for (auto &e : ev_pool) hipEventDestroy(e);
    hipEventDestroy(ev_next);
    hipEventDestroy(ev_panel);
    hipEventDestroy(ev_join);
    hipStreamSynchronize(stream2);
    hipStreamDestroy(stream2);
    hipStreamDestroy(own_stream);
}

template <typename T>
PanelWork<T> Handle::panel_work()
{
    PanelWork<T> pw;
    pw.mail = panel_mail;
    pw.mail_state = &panel_mail_state;
    pw.ctr = ctr;
    pw.abort_flag = abort_flag;
    pw.zero_col = zero_col;
    pw.spin_limit = spin_limit;
    pw.target_wgs = panel_wgs;
    return pw;
}

TrsvWork Handle::trsv_work(int n)
{
    TrsvWork tw;
    // 16 right-hand sides x 2 words per row, rows rounded up to the larger
    // block height of the two solvers (64 / 128)
    const size_t words = (size_t)ceil_div(std::max(n, 1), 128) * 128 * 16 * 2;
    const bool grow = words * sizeof(uint64_t) > bufs.at(BUF_TRSV_MAIL).cap;
    tw.mail = static_cast<uint64_t *>(buffer(BUF_TRSV_MAIL, words * sizeof(uint64_t)));
    if (grow) {
        // epoch 0 everywhere: no word of a fresh mailbox matches a launch
        fill_u32(reinterpret_cast<unsigned *>(tw.mail), (int)(words * 2), 0u, stream);
        SLA_HIP(hipStreamSynchronize(stream));
    }
    tw.mail_words = words;
    tw.ticket = ticket;
    tw.epoch = trsv_epoch;
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

void *Handle::buffer(int id, size_t bytes)
{
    Buf &b = bufs.at(id);
    if (bytes > b.cap) {
        SLA_HIP(hipStreamSynchronize(stream));
        SLA_HIP(hipStreamSynchronize(stream2));
        if (b.p) SLA_HIP(hipFree(b.p));
        b.p = alloc_dev<char>(bytes);
        b.cap = bytes;
    }
    return b.p;
}

// --- per-kernel timing -----------------------------------------------------
void Handle::timing_begin(int cat, hipStream_t s)
{
    if (!profile) return;
    if (!s) s = stream;
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
    SLA_HIP(hipEventRecord(e0, s));
    pending.push_back({cat, 0.0, e0, e1, s});
}

void Handle::timing_end(double work_units, hipStream_t s)
{
    if (!profile) return;
    if (!s) s = stream;
    // the open interval of this stream
    for (auto it = pending.rbegin(); it != pending.rend(); ++it) {
        if (it->s == s && !it->closed) {
            it->work = work_units;
            it->closed = true;
            SLA_HIP(hipEventRecord(it->e1, s));
            return;
        }
    }
}

unsigned long long *Handle::next_kstamp(double w)
{
    if (!profile || kstamp_used >= KSTAMP_CAP) return nullptr;
    kstamp_work.push_back(w);
    return kstamps + 2 * kstamp_used++;
}

void Handle::timing_collect()
{
    if (!profile) return;
    SLA_HIP(hipStreamSynchronize(stream));
    SLA_HIP(hipStreamSynchronize(stream2));
    if (kstamp_used > 0) {
        std::vector<unsigned long long> t(2 * (size_t)kstamp_used);
        SLA_HIP(hipMemcpy(t.data(), kstamps, t.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost));
        for (int i = 0; i < kstamp_used; i++) {
            if (t[2 * i + 1] < t[2 * i]) continue;  // the stamped kernel did not run
            exec.ms += (double)(t[2 * i + 1] - t[2 * i]) / wall_khz;
            exec.work += kstamp_work[i];
            exec.launches += 1;
        }
        reset_kstamps(kstamps, kstamp_used, stream);
        SLA_HIP(hipStreamSynchronize(stream));
        kstamp_used = 0;
        kstamp_work.clear();
    }
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
    exec = Stat{};
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
    int schur = SCHUR_NATIVE;  // precision of the outer trailing update
    bool l_packed = false;     // L21 of the current block column is packed
    hipStream_t s_main = nullptr, s_panel = nullptr;  // update stream, panel stream
    const T *inv_of[2] = {nullptr, nullptr};  // L11 whose inverses each stream's area holds
    int inv_m[2] = {0, 0};

    T *at(int i, int j) const { return A + (size_t)i + (size_t)j * lda; }

    // Pair list buffer q (two alternate between consecutive block columns).
    int *pairs(int q) const { return h.pairs + q * 2048; }
    int *pair_count(int q) const { return h.pair_counts + q; }

    void make_pairs(int k1, int k2, int q, hipStream_t s)
    {
        h.timing_begin(CAT_LASWP, s);
        laswp_pairs(ipiv, k1, k2, M, pairs(q), pair_count(q), h.bad_pivot, s);
        h.timing_end(0.0, s);
    }

    void swap_block(int k1, int k2, int q, int c0, int ncols, hipStream_t s)
    {
        if (ncols <= 0) return;
        h.timing_begin(CAT_LASWP, s);
        laswp_apply<T>(at(0, c0), lda, 0, ncols, pairs(q), pair_count(q), 2 * (k2 - k1), s);
        h.timing_end(32.0 * (double)(k2 - k1) * ncols, s);
    }

    void gemm(int m, int n, int k, const T *a, const T *b, T *c, hipStream_t s);

    // One panel leaf; returns the device-wide barrier arrivals it adds.
    unsigned leaf(T *a, int m, int w, int lcols, int rcols, int row0, int col0, hipStream_t s);

    // Outer trailing update A22 -= L21 U12 of one block column (k = its
    // width): the storage-precision GEMM, or for FP32 storage the BF16 / FP8
    // MFMA update on operands packed once per block column.
    void update(int m, int n, int k, const T *l21, const T *u12, T *c, hipStream_t s);

    // B[m x n] <- L^{-1} B for the unit lower m x m L at l.  With trsm_mfma
    // (default) every block of 33 .. 512 rows is one blocked launch (32-row
    // diagonal blocks by their inverses, the products on MFMA); otherwise
    // recursive halving down to <= 32-row diagonal solves (or one strip
    // launch for a small block inside the panel), the off-diagonal products
    // on the GEMMs (the 32-row blocking mirrors PB_CptrsmAB.c:359-412).
    void trsm(int m, int n, const T *l, T *b, hipStream_t s)
    {
        if (m <= 0 || n <= 0) return;
        if (h.trsm_mfma && m > 32 && m <= 512) {
            // one inverse area per stream (the panel stream's solves overlap
            // the main stream's)
            T *inv = static_cast<T *>(h.buffer(BUF_TRSM_INV, 2 * trsm_inv_elems(512) * sizeof(T)));
            const int slot = s != s_main;
            if (slot) inv += trsm_inv_elems(512);
            // the block row's two solves (next block column, then the rest)
            // share L11: the second reuses the inverses
            const bool ready = inv_of[slot] == l && inv_m[slot] == m;
            inv_of[slot] = l;
            inv_m[slot] = m;
            h.timing_begin(CAT_TRSM, s);
            trsm_llnu_blocked<T>(m, n, l, lda, b, lda, inv, ready, s);
            h.timing_end((double)m * m * n, s);
            return;
        }
        // <= 32 rows, or a small block inside the panel: one strip launch;
        // for FP32 storage also up to 1024 columns (the next block column's
        // U12 then costs 3 launches instead of 31 -- it is on the critical
        // path; the FP64 path keeps the rounding order its pivots were
        // checked with)
        const int wide = std::is_same<T, float>::value ? 1024 : 256;
        if (m <= 32 || (m <= 256 && n <= wide)) {
            h.timing_begin(CAT_TRSM, s);
            trsm_llnu<T>(m, n, l, lda, b, lda, s);
            h.timing_end((double)m * m * n, s);
            return;
        }
        const int m1 = ((m / 2 + 31) / 32) * 32;
        trsm(m1, n, l, b, s);
        gemm(m - m1, n, m1, l + m1, b, b + m1, s);
        trsm(m - m1, n, l + (size_t)m1 + (size_t)m1 * lda, b + m1, s);
    }

    // Recursive panel on rows r0.., columns c0 .. c0+n-1 (global indices) of
    // the block column [pc0, pc1).  Every leaf applies its interchanges to
    // the whole block column, so no interchange step is needed in between.
    void panel(int r0, int c0, int n, int pc0, int pc1, hipStream_t s)
    {
        const int m = M - r0;
        if (n <= h.leaf_w) {
            h.timing_begin(CAT_PANEL, s);
            h.ctr_base += leaf(at(r0, c0), m, n, c0 - pc0, pc1 - c0 - n, r0, c0, s);
            h.timing_end((double)m * n * n, s);
            return;
        }
        int n1 = (n / 2 / h.leaf_w) * h.leaf_w;
        if (n1 < h.leaf_w) n1 = h.leaf_w;
        panel(r0, c0, n1, pc0, pc1, s);
        trsm(n1, n - n1, at(r0, c0), at(r0, c0 + n1), s);
        gemm(m - n1, n - n1, n1, at(r0 + n1, c0), at(r0, c0 + n1), at(r0 + n1, c0 + n1), s);
        panel(r0 + n1, c0 + n1, n - n1, pc0, pc1, s);
    }

    // Right-looking blocked LU (SRC/pdgetrf.f:227-295) with a look-ahead of
    // one block column.  Main stream S0, step of block column p = [j, jn):
    //   apply p's interchanges to the next block column and to the kept row
    //   permutation;
    //   U12 and the Schur update for the NEXT block column only -> event;
    //   apply p's interchanges to the columns left of p (nothing but these
    //   interchanges touches them any more) and to the rest of the trailing
    //   columns; U12 and the Schur update for the rest.
    // Panel stream S1, once the next block column is updated: factor it and
    // build its pair list.  S0 waits for S1 before its next step.  Every
    // interchange runs on S0, so those of one column range always arrive in
    // panel order, and the panel stream -- the critical path whenever the
    // remaining trailing matrix is small -- carries only the panel.
    int run()
    {
        const int mn = std::min(M, N), nb = h.nb;
        if (nb > 512) throw std::logic_error("getrf: NB above 512");
        hipStream_t s0 = s_main, s1 = s_panel;
        {
            const int jb = std::min(nb, mn);
            panel(0, 0, jb, 0, jb, s0);
            make_pairs(0, jb, 0, s0);
        }
        for (int j = 0, p = 0; j < mn; j += nb, p++) {
            const int jb = std::min(nb, mn - j), jn = j + jb, q = p & 1;
            const bool has_next = jn < mn;
            const int jb2 = has_next ? std::min(nb, mn - jn) : 0;
            // block p's interchanges: before the hand-off only on the next
            // block column, so that the next panel starts as early as
            // possible; on the columns left of p and on the rest of the
            // trailing columns after it (the panel stream touches neither)
            swap_block(j, jn, q, jn, std::min(jb2, N - jn), s0);
            laswp_apply<int>(h.perm, M, 0, 1, pairs(q), pair_count(q), 2 * jb, s0);
            if (jn >= N) {
                swap_block(j, jn, q, 0, j, s0);
                break;
            }
            l_packed = false;
            if (has_next) {
                trsm(jb, jb2, at(j, j), at(j, jn), s0);
                update(M - jn, jb2, jb, at(jn, j), at(j, jn), at(jn, jn), s0);
                SLA_HIP(hipEventRecord(h.ev_next, s0));
                SLA_HIP(hipStreamWaitEvent(s1, h.ev_next, 0));
                panel(jn, jn, jb2, jn, jn + jb2, s1);
                make_pairs(jn, jn + jb2, q ^ 1, s1);
                SLA_HIP(hipEventRecord(h.ev_panel, s1));
            }
            const int c = jn + jb2, rest = N - c;
            swap_block(j, jn, q, 0, j, s0);
            swap_block(j, jn, q, c, rest, s0);
            if (rest > 0) {
                trsm(jb, rest, at(j, j), at(j, c), s0);
                if (jn < M) update(M - jn, rest, jb, at(jn, j), at(j, c), at(jn, c), s0);
            }
            if (has_next) SLA_HIP(hipStreamWaitEvent(s0, h.ev_panel, 0));
        }
        return 0;
    }
};

template <>
void Lu<double>::gemm(int m, int n, int k, const double *a, const double *b, double *c,
                      hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    h.timing_begin(CAT_GEMM_PANEL, s);
    dgemm_nn_sub(m, n, k, a, lda, b, lda, c, lda, s);
    h.timing_end(2.0 * m * n * k, s);
}

template <>
void Lu<double>::update(int m, int n, int k, const double *l21, const double *u12, double *c,
                        hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    h.timing_begin(CAT_GEMM, s);
    unsigned long long *st = h.next_kstamp(2.0 * (m / 128 * 128) * (n / 128 * 128) * k);
    dgemm_nn_sub(m, n, k, l21, lda, u12, lda, c, lda, s, true, h.update_cus, st);
    h.timing_end(2.0 * m * n * k, s);
}

template <>
unsigned Lu<double>::leaf(double *a, int m, int w, int lcols, int rcols, int row0, int col0,
                          hipStream_t s)
{
    return panel_leaf<double>(a, lda, m, w, lcols, rcols, row0, col0, ipiv, pw, s);
}

template <>
unsigned Lu<float>::leaf(float *a, int m, int w, int lcols, int rcols, int row0, int col0,
                         hipStream_t s)
{
    if (h.pivoting == 1 && panel_tslu_fits<float>(m, w, pw.target_wgs)) {
        char *b = static_cast<char *>(h.buffer(BUF_TSLU, tslu_slot_bytes(sizeof(float))));
        TsluSlots<float> slots;
        slots.cand_r = reinterpret_cast<int *>(b);
        slots.cand_v = reinterpret_cast<float *>(b + 32 * 32 * 4);
        slots.top_v = slots.cand_v + 32 * 32 * 32;
        return panel_tslu<float>(a, lda, m, w, lcols, rcols, row0, col0, ipiv, pw, slots,
                                 h.ctr_base, s);
    }
    return panel_leaf<float>(a, lda, m, w, lcols, rcols, row0, col0, ipiv, pw, s);
}

template <>
void Lu<float>::gemm(int m, int n, int k, const float *a, const float *b, float *c, hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    h.timing_begin(CAT_GEMM_PANEL, s);
    sgemm_nn_sub(m, n, k, a, lda, b, lda, c, lda, s);
    h.timing_end(2.0 * m * n * k, s);
}

template <>
void Lu<float>::update(int m, int n, int k, const float *l21, const float *u12, float *c,
                       hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (schur == SCHUR_NATIVE) {
        h.timing_begin(CAT_GEMM, s);
        unsigned long long *st = h.next_kstamp(2.0 * (m / 128 * 128) * (n / 128 * 128) * k);
        sgemm_nn_sub(m, n, k, l21, lda, u12, lda, c, lda, s, true, 2 * h.update_cus,
                     h.update_lds_kb * 1024, st);
        h.timing_end(2.0 * m * n * k, s);
        return;
    }
    if (const int split = split_of_schur(schur)) {
        // split 16-bit products: S planes of each operand, k padded to 32
        const int S = split_slices(split);
        const int kpad = ceil_div(k, 32) * 32;
        const size_t lplane = (size_t)m * kpad, uplane = (size_t)n * kpad;
        char *lp = static_cast<char *>(h.buffer(BUF_PACK_L, (size_t)M * kpad * 2 * S + 256));
        char *up = static_cast<char *>(h.buffer(BUF_PACK_U, (size_t)N * kpad * 2 * S + 256));
        float *absmax = static_cast<float *>(h.buffer(BUF_SCALES, 64 * sizeof(float)));
        const bool scaled = split == SPLIT_FP16X3;
        h.timing_begin(CAT_OTHER, s);
        if (!l_packed) {
            if (scaled) {
                fill_u32(reinterpret_cast<unsigned *>(absmax), 1, 0u, s);
                absmax_scale(m, k, l21, lda, absmax, s);
            }
            pack_split(split, true, m, k, kpad, l21, lda, scaled ? absmax : nullptr, lp, lplane, s);
        }
        if (scaled) {
            fill_u32(reinterpret_cast<unsigned *>(absmax) + 1, 1, 0u, s);
            absmax_scale(k, n, u12, lda, absmax + 1, s);
        }
        pack_split(split, false, n, k, kpad, u12, lda, scaled ? absmax + 1 : nullptr, up, uplane,
                   s);
        if (!scaled) fill_u32(reinterpret_cast<unsigned *>(absmax), 2, 0u, s);
        h.timing_end(0.0, s);
        l_packed = true;
        h.timing_begin(CAT_GEMM, s);
        unsigned long long *st = h.next_kstamp(2.0 * (m / 128 * 128) * (n / 128 * 128) * k);
        gemm_split_nn_sub(split, m, n, kpad, lp, kpad, lplane, up, kpad, uplane, absmax, c, lda, s,
                          st);
        h.timing_end(2.0 * m * n * k, s);
        return;
    }
    // operands k-contiguous, k padded with zeros to the MFMA's K step
    const int kq = schur == SCHUR_BF16 ? 32 : 128;
    const int kpad = ceil_div(k, kq) * kq;
    const size_t esz = schur == SCHUR_BF16 ? 2 : 1;
    char *lp = static_cast<char *>(h.buffer(BUF_PACK_L, (size_t)M * kpad * esz));
    char *up = static_cast<char *>(h.buffer(BUF_PACK_U, (size_t)N * kpad * esz));
    float *scales = static_cast<float *>(h.buffer(BUF_SCALES, 64 * sizeof(float)));
    h.timing_begin(CAT_OTHER, s);
    if (schur == SCHUR_BF16) {
        if (!l_packed) pack_rows_bf16(m, k, kpad, l21, lda, reinterpret_cast<__bf16 *>(lp), s);
        pack_cols_bf16(k, kpad, n, u12, lda, reinterpret_cast<__bf16 *>(up), s);
    } else {
        if (!l_packed) {
            fill_u32(reinterpret_cast<unsigned *>(scales), 1, 0u, s);
            absmax_scale(m, k, l21, lda, scales, s);
            pack_rows_fp8(m, k, kpad, l21, lda, scales, reinterpret_cast<uint8_t *>(lp), s);
        }
        fill_u32(reinterpret_cast<unsigned *>(scales) + 1, 1, 0u, s);
        absmax_scale(k, n, u12, lda, scales + 1, s);
        pack_cols_fp8(k, kpad, n, u12, lda, scales + 1, reinterpret_cast<uint8_t *>(up), s);
    }
    h.timing_end(0.0, s);
    l_packed = true;
    h.timing_begin(CAT_GEMM, s);
    unsigned long long *st = h.next_kstamp(2.0 * (m / 128 * 128) * (n / 128 * 128) * k);
    if (schur == SCHUR_BF16)
        gemm_bf16_nn_sub(m, n, kpad, reinterpret_cast<const __bf16 *>(lp), kpad,
                         reinterpret_cast<const __bf16 *>(up), kpad, c, lda, s, st);
    else
        gemm_fp8_nn_sub(m, n, kpad, reinterpret_cast<const uint8_t *>(lp), kpad,
                        reinterpret_cast<const uint8_t *>(up), kpad, scales, c, lda, s, st);
    h.timing_end(2.0 * m * n * k, s);
}

}  // namespace

template <typename T>
static int getrf_device(Handle &h, int m, int n, T *dA, int lda, int *dipiv, int schur)
{
    if (m <= 0 || n <= 0) return 0;
    const int mn = std::min(m, n);
    h.ensure_perm(m);
    if (h.trsm_mfma) h.buffer(BUF_TRSM_INV, 2 * trsm_inv_elems(512) * sizeof(T));
    h.ctr_base = 0;
    fill_int(h.ctl, 6, 0, h.stream);         // counter, flags
    fill_int(h.zero_col, 1, INT_MAX, h.stream);
    iota_int(h.perm, m, 0, h.stream);
    // streams of the factorization: the caller's (update) + the handle's second
    // one (panel); the panel stream starts after everything queued so far on
    // the caller's stream, which waits for it at the end
    hipStream_t s0 = h.stream, s1 = h.stream2;
    if (!h.lookahead) s1 = s0;
    SLA_HIP(hipEventRecord(h.ev_join, h.stream));
    if (s0 != h.stream) SLA_HIP(hipStreamWaitEvent(s0, h.ev_join, 0));
    if (s1 != h.stream) SLA_HIP(hipStreamWaitEvent(s1, h.ev_join, 0));
    Lu<T> lu{h, dA, lda, m, n, dipiv, h.panel_work<T>()};
    lu.schur = schur;
    lu.s_main = s0;
    lu.s_panel = s1;
    lu.run();
    for (hipStream_t sx : {s0, s1}) {
        if (sx == h.stream) continue;
        SLA_HIP(hipEventRecord(h.ev_join, sx));
        SLA_HIP(hipStreamWaitEvent(h.stream, h.ev_join, 0));
    }
    int flags[3];
    SLA_HIP(hipMemcpyAsync(flags, h.ctl, sizeof flags, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    h.timing_collect();
    if (flags[1]) throw std::runtime_error("getrf: a device-wide barrier timed out (panel)");
    h.perm_key_a = dA;
    h.perm_key_ipiv = dipiv;
    h.perm_m = m;
    const int zc = flags[2];
    return (zc != INT_MAX && zc <= mn) ? zc : 0;
}

int dgetrf_device(Handle &h, int m, int n, double *dA, int lda, int *dipiv)
{
    return getrf_device<double>(h, m, n, dA, lda, dipiv, SCHUR_NATIVE);
}

int sgetrf_device(Handle &h, int m, int n, float *dA, int lda, int *dipiv, int schur)
{
    if (schur < SCHUR_NATIVE || schur > SCHUR_FP16X3)
        throw std::invalid_argument("sgetrf: unknown Schur-update precision");
    return getrf_device<float>(h, m, n, dA, lda, dipiv, schur);
}

// Row permutation of the factorization from IPIV alone (for a solve that did
// not follow this handle's own factorization).
void perm_from_ipiv(Handle &h, int n, const int *dipiv)
{
    h.perm_key_a = nullptr;
    h.perm_key_ipiv = nullptr;
    h.perm_m = 0;
    h.ensure_perm(n);
    iota_int(h.perm, n, 0, h.stream);
    for (int k = 0; k < n; k += 512) {
        const int k2 = std::min(n, k + 512);
        laswp_pairs(dipiv, k, k2, n, h.pairs, h.pair_count, h.bad_pivot, h.stream);
        laswp_apply<int>(h.perm, n, 0, 1, h.pairs, h.pair_count, 2 * (k2 - k), h.stream);
    }
}

void getrs_perm_device(Handle &h, bool trans, int n, int nrhs, const double *dA, int lda,
                       double *dB, int ldb)
{
    if (n <= 0 || nrhs <= 0) return;
    const int chunk = std::min(nrhs, 16);
    h.ensure_work((size_t)n * chunk * sizeof(double));
    double *W = reinterpret_cast<double *>(h.work);
    TrsvWork tw = h.trsv_work(n);
    for (int j = 0; j < nrhs; j += 16) {
        const int nr = std::min(16, nrhs - j);
        double *Bj = dB + (size_t)j * ldb;
        if (!trans) {
            // P B, then L^{-1}, then U^{-1} (pdgetrs.f:255-266)
            h.timing_begin(CAT_LASWP);
            gather_rows<double>(Bj, ldb, W, n, h.perm, n, nr, h.stream);
            h.timing_end(16.0 * n * nr);
            h.timing_begin(CAT_TRSV);
            trsv_persistent<double, double>(TrsvOp::L, n, nr, dA, lda, W, n, tw, h.stream);
            h.timing_end(4.0 * (double)n * n);
            h.timing_begin(CAT_TRSV);
            trsv_persistent<double, double>(TrsvOp::U, n, nr, dA, lda, W, n, tw, h.stream);
            h.timing_end(4.0 * (double)n * n);
            SLA_HIP(hipMemcpy2DAsync(Bj, (size_t)ldb * sizeof(double), W, (size_t)n * sizeof(double),
                                     (size_t)n * sizeof(double), nr, hipMemcpyDeviceToDevice,
                                     h.stream));
        } else {
            // U^{-T}, then L^{-T}, then P^T (pdgetrs.f:273-284)
            SLA_HIP(hipMemcpy2DAsync(W, (size_t)n * sizeof(double), Bj, (size_t)ldb * sizeof(double),
                                     (size_t)n * sizeof(double), nr, hipMemcpyDeviceToDevice,
                                     h.stream));
            h.timing_begin(CAT_TRSV);
            trsv_persistent<double, double>(TrsvOp::UT, n, nr, dA, lda, W, n, tw, h.stream);
            h.timing_end(4.0 * (double)n * n);
            h.timing_begin(CAT_TRSV);
            trsv_persistent<double, double>(TrsvOp::LT, n, nr, dA, lda, W, n, tw, h.stream);
            h.timing_end(4.0 * (double)n * n);
            h.timing_begin(CAT_LASWP);
            scatter_rows<double>(W, n, Bj, ldb, h.perm, n, nr, h.stream);
            h.timing_end(16.0 * n * nr);
        }
    }
    h.trsv_epoch = tw.epoch;
}

void check_solve_flags(Handle &h, const char *routine)
{
    int flags[6];
    SLA_HIP(hipMemcpyAsync(flags, h.ctl, sizeof flags, hipMemcpyDeviceToHost, h.stream));
    SLA_HIP(hipStreamSynchronize(h.stream));
    h.timing_collect();
    if (flags[5]) throw std::runtime_error(std::string(routine) + ": IPIV holds a row outside 1..N");
    if (flags[1])
        throw std::runtime_error(std::string(routine) +
                                 ": a dependency wait timed out (triangular solve)");
}

int dgetrs_device(Handle &h, int n, int nrhs, const double *dA, int lda, const int *dipiv,
                  double *dB, int ldb, bool trans)
{
    if (n <= 0 || nrhs <= 0) return 0;
    fill_int(h.bad_pivot, 1, 0, h.stream);
    fill_int(h.abort_flag, 1, 0, h.stream);
    if (!(h.perm_key_a == dA && h.perm_key_ipiv == dipiv && h.perm_m == n))
        perm_from_ipiv(h, n, dipiv);
    getrs_perm_device(h, trans, n, nrhs, dA, lda, dB, ldb);
    check_solve_flags(h, "dgetrs");
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
