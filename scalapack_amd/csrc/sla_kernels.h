// sla_kernels.h -- host-side launchers of the HIP kernels (one stream each).
#pragma once

#include "sla_common.h"

namespace sla {

// ---- Schur-complement update: C[m x n] -= A[m x k] * B[k x n] (FP64) -----
// trailing: the launch is PDGETRF's outer trailing update (a kernel symbol of
// its own in traces); max_wgs > 0 caps its workgroups (a persistent grid that
// leaves CUs to the panel running beside it).
// stamp: launch_stamp pair of the interior (128 x 128-tiled) launch of a
// trailing update, or null.
void dgemm_nn_sub(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                  double *C, int ldc, hipStream_t s, bool trailing = false, int max_wgs = 0,
                  unsigned long long *stamp = nullptr);

// ---- thin GEMMs of the recursive panel and TRSM (gemm_small.hip) -----------
// C -= A * B when n, k <= 64 with m >= 4n (tall-narrow) or m, k <= 64
// (short-wide); returns false (nothing launched) for other shapes.
constexpr int SMALL_NMAX = 64, SMALL_KMAX = 64;
template <typename T>
bool gemm_small(int m, int n, int k, const T *A, int lda, const T *B, int ldb, T *C, int ldc,
                hipStream_t s);

// ---- low-precision Schur-complement updates (FP32 storage) -----------------
// C[m x n] -= A[m x k] * B[k x n], FP32 operands on FP32 MFMA.
// reserve_lds: bytes of LDS reserved per workgroup of the trailing launch
// (limits the workgroups sharing a CU; the kernel itself does not use it).
void sgemm_nn_sub(int m, int n, int k, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, hipStream_t s, bool trailing = false, int max_wgs = 0,
                  int reserve_lds = 0, unsigned long long *stamp = nullptr);
// Same with BF16 operands packed k-contiguous: Ap row-major m x k (ld ldak),
// Bp column-major k x n (ld ldbk); FP32 accumulation into C.
void gemm_bf16_nn_sub(int m, int n, int k, const __bf16 *Ap, int ldak, const __bf16 *Bp, int ldbk,
                      float *C, int ldc, hipStream_t s, unsigned long long *stamp = nullptr);
// FP8-E4M3 (OCP) operands packed like the BF16 ones; the products are scaled
// back by sa * sb (the per-panel scales the operands were divided by).
void gemm_fp8_nn_sub(int m, int n, int k, const uint8_t *Ap, int ldak, const uint8_t *Bp,
                     int ldbk, const float *scales, float *C, int ldc, hipStream_t s,
                     unsigned long long *stamp = nullptr);
// Packing: A (m x k, column-major) -> row-major with k padded by zeros to
// kpad; B (k x n, column-major) -> column-major with ld kpad, zero padded.
void pack_rows_bf16(int m, int k, int kpad, const float *A, int lda, __bf16 *P, hipStream_t s);
void pack_cols_bf16(int k, int kpad, int n, const float *B, int ldb, __bf16 *P, hipStream_t s);
// FP8 packing divides by the scale held at *scale (written by absmax_scale).
void pack_rows_fp8(int m, int k, int kpad, const float *A, int lda, const float *scale,
                   uint8_t *P, hipStream_t s);
void pack_cols_fp8(int k, int kpad, int n, const float *B, int ldb, const float *scale,
                   uint8_t *P, hipStream_t s);
// *scale = max|A| over an m x n block (the FP8 and FP16 packings derive
// their block scales from it); *scale must hold 0 on entry (bit-pattern max).
void absmax_scale(int m, int n, const float *A, int lda, float *scale, hipStream_t s);

// Split 16-bit products (gemm_lowp.hip): an FP32 operand as S 16-bit slices,
// the slice products of order below S on 16-bit MFMA, FP32 accumulation.
enum Split {
    SPLIT_BF16X3 = 1,  // 2 BF16 slices, 3 products, ~2^-16 |a||b|
    SPLIT_BF16X6 = 2,  // 3 BF16 slices, 6 products, ~2^-24 |a||b|
    SPLIT_FP16X3 = 3   // 2 FP16 slices (power-of-two block scale), 3 products, ~2^-21 |a||b|
};
inline int split_slices(int split) { return split == SPLIT_BF16X6 ? 3 : 2; }
// Packs A (rows: m x k -> S row-major planes) or B (!rows: k x n -> S
// column-major planes), k padded with zeros to kpad, planes `plane` elements
// apart; absmax -> max|block| (absmax_scale), read by the FP16 split only.
void pack_split(int split, bool rows, int d0, int k, int kpad, const float *A, int lda,
                const float *absmax, void *P, size_t plane, hipStream_t s);
// C[m x n] -= A B from packed planes; absmax[0], absmax[1] the maxima the
// A and B packings were scaled by (FP16 split).
void gemm_split_nn_sub(int split, int m, int n, int k, const void *Ap, int ldak, size_t aplane,
                       const void *Bp, int ldbk, size_t bplane, const float *absmax, float *C,
                       int ldc, hipStream_t s, unsigned long long *stamp = nullptr);

// ---- precision conversions -------------------------------------------------
// B = (float)A; sets *overflow when some |A(i,j)| exceeds the FP32 range.
void convert_f64_f32(int m, int n, const double *A, int lda, float *B, int ldb, int *overflow,
                     hipStream_t s);

// ---- panel leaf (PDGETF2 on <= 32 columns) --------------------------------
// Mailbox of the leaf's per-column exchange: 2 regions x 32 column blocks of
// PANEL_MAIL_COL 64-bit words -- PANEL_MAIL_GMAX candidates of up to 33 words
// (row index, 32 values), then the 32 words of row j at PANEL_MAIL_ROWJ.
// Every word holds the empty pattern (all ones) until a launch writes it.
constexpr int PANEL_MAIL_GMAX = 128;
constexpr int PANEL_MAIL_ROWJ = PANEL_MAIL_GMAX * 33;
constexpr int PANEL_MAIL_COL = PANEL_MAIL_ROWJ + 32;
constexpr size_t PANEL_MAIL_REGION = 32 * (size_t)PANEL_MAIL_COL;
constexpr size_t PANEL_MAIL_WORDS = 2 * PANEL_MAIL_REGION;

// Host-side record of the mailbox's last launch (whose words the next
// launch empties) and the launch count (whose parity picks the region).
struct PanelMailState {
    unsigned launches = 0;
    int prev_wa = 0, prev_g = 0, prev_cs = 0;
};

template <typename T>
struct PanelWork {
    uint64_t *mail;             // PANEL_MAIL_WORDS words, see above
    PanelMailState *mail_state; // host record shared by every leaf of the handle
    unsigned *ctr;       // monotonic arrival counter of the device-wide barrier
    int *abort_flag;     // set by a wait that lasted past spin_limit
    int *zero_col;       // min over zero pivots of the 1-based column (INT_MAX: none)
    uint64_t spin_limit; // wall-clock ticks a barrier may wait
    int target_wgs;      // preferred number of workgroups per leaf launch
};

// Factors the m x wa leaf at A (global rows row0.., columns col0.. relative to
// the factored matrix) in place; writes ipiv[row0 .. row0+wa-1] as 1-based
// global rows and applies the interchanges also to the lcols columns left
// and rcols columns right of the leaf (the rest of its block column).
// Launches on one handle are ordered (one stream at a time).  Returns the
// barrier arrivals the launch adds to *ctr: none.
template <typename T>
unsigned panel_leaf(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                    int *ipiv, const PanelWork<T> &pw, hipStream_t s);

// ---- panel leaf with tournament pivoting (low-precision factorization) -------
template <typename T>
struct TsluSlots {
    int *cand_r;  // [32 workgroups][32] candidate rows (leaf-relative)
    T *cand_v;    // [32][32 * 32] their original values, column by column
    T *top_v;     // [32][32] the leaf's top rows, original values
};
constexpr size_t tslu_slot_bytes(size_t esz) { return 32 * 32 * 4 + (32 * 32 * 32 + 32 * 32) * esz + 256; }
// Whether the tournament leaf handles an m x wa leaf (wa <= 32 <= m, m at
// most 32768 rows).
template <typename T>
bool panel_tslu_fits(int m, int wa, int target_wgs);
// Factors the m x wa leaf like panel_leaf, choosing the pivots by a
// tournament (one device-wide barrier); returns the barrier arrivals.
template <typename T>
unsigned panel_tslu(T *A, int lda, int m, int wa, int lcols, int rcols, int row0, int col0,
                    int *ipiv, const PanelWork<T> &pw, const TsluSlots<T> &slots, unsigned base,
                    hipStream_t s);

// ---- row interchanges (PDLASWP) --------------------------------------------
// Net effect of the forward interchanges ipiv[k1 .. k2-1] (0-based global row
// indices of the pivoted rows, ipiv 1-based global) as a list of
// (dst, src) global rows: after the interchanges row dst holds what row src
// held before.  At most 2*(k2-k1) pairs; k2-k1 <= 512.  A pivot row outside
// [1, nrows] sets *bad and yields an empty list.
void laswp_pairs(const int *ipiv, int k1, int k2, int nrows, int *pairs, int *count, int *bad,
                 hipStream_t s);
// Applies a pair list to columns [0, ncols) of A whose row 0 is global row
// `roff`; `inverse` applies the inverse permutation (the same interchanges
// replayed backward).
template <typename T>
void laswp_apply(T *A, int lda, int roff, int ncols, const int *pairs, const int *count,
                 int max_pairs, hipStream_t s, bool inverse = false);
// Column interchanges i <-> ipiv[i]-1, i = k1 .. k2-1, over rows [0, nrows).
void laswp_cols(double *A, int lda, int nrows, int k1, int k2, const int *ipiv, bool forward,
                hipStream_t s);

// ---- triangular solves -----------------------------------------------------
// B[m x n] <- L^{-1} B, L unit lower triangular m x m, m <= 256
// (PDTRSM 'L','L','N','U').
template <typename T>
void trsm_llnu(int m, int n, const T *L, int ldl, T *B, int ldb, hipStream_t s);
// The same for m <= 512 with the products on MFMA (inverses of the 32 x 32
// diagonal blocks, then block substitution); inv: trsm_inv_elems(m) of T,
// already holding the inverses of this L when inv_ready.
constexpr size_t trsm_inv_elems(int m) { return (size_t)((m + 31) / 32) * 1024; }
template <typename T>
void trsm_llnu_blocked(int m, int n, const T *L, int ldl, T *B, int ldb, T *inv, bool inv_ready,
                       hipStream_t s);
// One launch per triangle for n x nrhs right-hand sides (nrhs <= 16), with L
// (unit lower) and U (upper, non-unit) stored together as PDGETRF leaves them:
// X <- L^{-1} X, U^{-1} X, U^{-T} X or L^{-T} X.
enum class TrsvOp { L, U, UT, LT };
struct TrsvWork {
    int *ticket;         // workgroup ticket counter (reset before each launch)
    uint64_t *mail;      // solved values of every block, tagged with the epoch
    size_t mail_words;   // capacity of mail (64 * 16 * 2 words per 64-row block)
    int epoch;           // host-side counter, one value per launch
    int *abort_flag;
    uint64_t spin_limit;
};
// TA: storage type of the triangle, T: type of X and of the arithmetic
// (<double, double>, <float, float>, <float, double>).
template <typename TA, typename T>
void trsv_persistent(TrsvOp op, int n, int nrhs, const TA *A, int lda, T *X, int ldx,
                     TrsvWork &tw, hipStream_t s);

// Block-inverse solves with FP32 factors (refinement corrections, GMRES-IR's
// preconditioner): trsv_block_inverses writes the FP64 inverses of the
// 128 x 128 diagonal blocks of L (plane 0) and U (plane 1) into inv
// (2 * ceil(n / 128) * 128 * 128 doubles); trsv_blockinv then solves
// X <- L^{-1} X (lower) or U^{-1} X with them, one launch per triangle.
constexpr int TRSV_INV_BS = 128;
inline size_t trsv_inverse_doubles(int n)
{
    return (size_t)2 * ((n + TRSV_INV_BS - 1) / TRSV_INV_BS) * TRSV_INV_BS * TRSV_INV_BS;
}
void trsv_block_inverses(int n, const float *A, int lda, double *inv, hipStream_t s);
template <typename T>
void trsv_blockinv(bool lower, int n, int nrhs, const float *A, int lda, const double *inv, T *X,
                   int ldx, TrsvWork &tw, hipStream_t s);

// ---- small helpers ---------------------------------------------------------
void fill_int(int *p, int n, int v, hipStream_t s);
// n launch-stamp pairs to (UINT64_MAX, 0): no start seen, no end seen
void reset_kstamps(unsigned long long *p, int n, hipStream_t s);
void fill_u32(unsigned *p, int n, unsigned v, hipStream_t s);
void iota_int(int *p, int n, int base, hipStream_t s);
// B[i, :] = Bsrc[perm[i], :] (row gather), n x nrhs
template <typename T>
void gather_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n, int nrhs,
                 hipStream_t s);
// B[perm[i], :] = Bsrc[i, :] (row scatter, the inverse permutation)
template <typename T>
void scatter_rows(const T *src, int lds, T *dst, int ldd, const int *perm, int n, int nrhs,
                  hipStream_t s);

}  // namespace sla
