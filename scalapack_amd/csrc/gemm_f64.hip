// gemm_f64.hip -- the trailing (Schur-complement) update of the LU
// factorization, A22 <- A22 - L21 * U12, as an FP64 MFMA GEMM for gfx950.
//
// This is the one local dgemm_ PDGEMM issues on a 1 x 1 grid
// (SRC/pdgetrf.f:288-290 -> PBLAS/SRC/pdgemm_.c:476 ->
//  PBLAS/SRC/PTOOLS/PB_CpgemmAB.c:333-352: dgemm_('N','N', M-J-JB+1,
//  N-J-JB+1, JB, -1, L21, LLD, U12, LLD, +1, A22, LLD)), and it also serves
// the updates inside the recursive panel and the block triangular solves.
//
// Tiling: a 256-thread workgroup (4 wave64s) owns a 128 x 128 tile of C and
// walks K in steps of 16 through double-buffered LDS tiles.  Each wave owns a
// 64 x 64 sub-tile as 4 x 4 accumulators of v_mfma_f64_16x16x4f64 (16 x 16
// outputs, 4 doubles per lane).  Lane layout of that instruction:
//   A operand : lane l holds A[l % 16][l / 16]         (16 rows x 4 k)
//   B operand : lane l holds B[l / 16][l % 16]         (4 k x 16 cols)
//   D         : lane l, element r holds D[l/16 + 4*r][l % 16]
// (checked on gfx950 with a one-hot probe), so for a fixed r the 64 lanes
// cover 4 consecutive rows of 16 columns: 32 contiguous bytes per column of
// column-major C.
#include "sla_common.h"
#include "sla_kernels.h"

#include <cstdlib>

namespace sla {

namespace {

constexpr int TM = 128, TN = 128, TK = 16;
constexpr int BPAD = 1;  // odd pad on the k-major B tile staggers LDS banks

typedef double d4_t __attribute__((ext_vector_type(4)));

// Maps the linear workgroup id to a (tile_m, tile_n) pair so that the
// workgroups one XCD receives (ids congruent mod 8 are dispatched to the same
// XCD) sweep a compact band of tiles that share L21 rows and U12 columns.
__device__ __forceinline__ void tile_coords(int id, int tiles_m, int tiles_n, int &tm, int &tn)
{
    const int total = tiles_m * tiles_n;
    int lid = id;
    if (total >= 64) {
        // Round-robin dispatch sends id to XCD id % 8; give each XCD a
        // contiguous range of the row-major tile order.
        const int per = total / 8, rem = total % 8;
        const int xcd = id % 8, k = id / 8;
        const int start = xcd * per + (xcd < rem ? xcd : rem);
        const int len = per + (xcd < rem ? 1 : 0);
        lid = (k < len) ? start + k : id;  // every id maps to a distinct tile
    }
    // Column-band order: consecutive ids walk down M inside a band of 8 N
    // tiles, so neighbours reuse the same U12 columns and L21 rows from L2.
    const int band = 8;
    const int bands_full = tiles_n / band;
    const int in_full = bands_full * band * tiles_m;
    if (lid < in_full) {
        const int b = lid / (band * tiles_m), r = lid % (band * tiles_m);
        tm = r % tiles_m;
        tn = b * band + r / tiles_m;
    } else {
        const int r = lid - in_full, w = tiles_n - bands_full * band;
        tm = r % tiles_m;
        tn = bands_full * band + r / tiles_m;
        (void)w;
    }
}

__global__ __launch_bounds__(256, 2) void dgemm_nn_sub_kernel(
    int m, int n, int k, const double *__restrict__ A, int lda,
    const double *__restrict__ B, int ldb, double *__restrict__ C, int ldc)
{
    __shared__ double As[2][TK][TM];
    __shared__ double Bs[2][TK][TN + BPAD];

    const int tiles_m = ceil_div(m, TM), tiles_n = ceil_div(n, TN);
    int tmi, tni;
    tile_coords(blockIdx.x, tiles_m, tiles_n, tmi, tni);
    const int bm = tmi * TM, bn = tni * TN;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;

    // Global -> register mapping.  A tile (128 x 16): lanes run down a column
    // of A (coalesced 512-byte rows per wave), 8 k per thread.
    const int a_row = tid & 127, a_k0 = (tid >> 7) * 8;
    // B tile (16 x 128): 16 lanes cover one column's 16 k (128 bytes).
    const int b_k = tid & 15, b_n0 = tid >> 4;  // columns b_n0 + 16*i

    const bool a_row_ok = (bm + a_row) < m;
    const double *Ap = A + (size_t)(bm + a_row);
    double ra[8], rb[8];

    auto load_tile = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int kk = k0 + a_k0 + i;
            ra[i] = (a_row_ok && kk < k) ? Ap[(size_t)kk * lda] : 0.0;
        }
        const int kk = k0 + b_k;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int nn = bn + b_n0 + 16 * i;
            rb[i] = (nn < n && kk < k) ? B[kk + (size_t)nn * ldb] : 0.0;
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 8; i++) As[buf][a_k0 + i][a_row] = ra[i];
#pragma unroll
        for (int i = 0; i < 8; i++) Bs[buf][b_k][b_n0 + 16 * i] = rb[i];
    };

    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

    const int nk = ceil_div(k, TK);
    load_tile(0);
    store_tile(0);
    __syncthreads();

    const int lr = lane & 15, lk = lane >> 4;
    for (int t = 0; t < nk; t++) {
        const int buf = t & 1;
        if (t + 1 < nk) load_tile((t + 1) * TK);
#pragma unroll
        for (int s = 0; s < TK / 4; s++) {
            double af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; i++) af[i] = As[buf][s * 4 + lk][wm + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; j++) bf[j] = Bs[buf][s * 4 + lk][wn + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (t + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }

    // Epilogue: C -= acc.  Lane owns rows lk, lk+4, lk+8, lk+12 of column lr
    // of each 16 x 16 block.
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int col = bn + wn + j * 16 + lr;
        if (col >= n) continue;
        double *Cc = C + (size_t)col * ldc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row0 = bm + wm + i * 16 + lk;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + 4 * r;
                if (row < m) Cc[row] = Cc[row] - acc[i][j][r];
            }
        }
    }
}

// The interior kernel: no LDS and no workgroup barrier.  Each wave owns a
// 64 x 64 tile of C and feeds v_mfma_f64_16x16x4f64 straight from HBM/L2 into
// VGPRs, three 16-deep K steps in flight.  Two index permutations make every
// lane's operands 32 contiguous bytes (one 32-byte load each):
//   rows : MFMA block i, operand row lr  <->  C row wm + 4*lr + i
//   K    : operand k-lane lk, step s     <->  k = k0 + 4*lk + s
// Both A and B use the same K permutation, so the sum is unchanged; the
// epilogue maps D element r of lane l (block row l/16 + 4r) back to C row
// wm + 4*(l/16 + 4r) + i, four consecutive rows for i = 0..3.
// Requires m, n multiples of 128, k a multiple of 16, 32-byte aligned column
// starts (lda, ldb, ldc multiples of 4 and aligned base pointers).
typedef double dv4_t __attribute__((ext_vector_type(4)));
constexpr int DEPTH = 3;

// TRAILING distinguishes the launches of the outer trailing update (1) from
// every other use (0) in kernel traces; the code is the same.
// The grid may be smaller than the tile count (a persistent launch that
// leaves CUs free for the concurrently running panel): workgroup b takes
// tiles b, b + grid, ...; grid is a multiple of 8, so a tile index keeps the
// XCD of the workgroup that processes it.
template <int TRAILING>
__global__ __launch_bounds__(256, 1) void dgemm_nn_sub_direct_kernel(
    int m, int n, int k, const double *__restrict__ A, int lda,
    const double *__restrict__ B, int ldb, double *__restrict__ C, int ldc,
    unsigned long long *stamp)
{
    launch_stamp_begin(stamp);
    const int tiles_m = m / 128, tiles_n = n / 128;
    for (int tile = blockIdx.x; tile < tiles_m * tiles_n; tile += gridDim.x) {
    int tmi, tni;
    tile_coords(tile, tiles_m, tiles_n, tmi, tni);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = tmi * 128 + (wave & 1) * 64, wn = tni * 128 + (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    const double *Ap = A + (wm + 4 * lr) + (size_t)(4 * lk) * lda;
    const double *Bp = B + (size_t)(wn + lr) * ldb + 4 * lk;
    const size_t bstep = (size_t)16 * ldb;

    dv4_t a[DEPTH][4], b[DEPTH][4];
    auto load = [&](int t, int q) {
        const size_t k0 = (size_t)16 * t;
#pragma unroll
        for (int s = 0; s < 4; s++) a[q][s] = *(const dv4_t *)(Ap + (k0 + s) * lda);
#pragma unroll
        for (int j = 0; j < 4; j++) b[q][j] = *(const dv4_t *)(Bp + j * bstep + k0);
    };
    d4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};

    const int nk = k / 16;
#pragma unroll
    for (int q = 0; q < DEPTH - 1; q++)
        if (q < nk) load(q, q);
    for (int t = 0; t < nk; t += DEPTH) {
#pragma unroll
        for (int q = 0; q < DEPTH; q++) {
            if (t + q < nk) {
                const int nt = t + q + DEPTH - 1;
                if (nt < nk) load(nt, (q + DEPTH - 1) % DEPTH);
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q][s][i], b[q][j][s],
                                                                            acc[i][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        double *Cc = C + (size_t)(wn + 16 * j + lr) * ldc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            dv4_t *cp = (dv4_t *)(Cc + wm + 4 * (lk + 4 * r));
            dv4_t c = *cp;
            c[0] -= acc[0][j][r];
            c[1] -= acc[1][j][r];
            c[2] -= acc[2][j][r];
            c[3] -= acc[3][j][r];
            *cp = c;
        }
    }
    }  // tile loop
    launch_stamp_end(stamp);
}

// The wide interior kernel: a 128 x 256 tile of C per 512-thread workgroup
// (one per CU, two waves per SIMD), each of the 8 waves a 64 x 64 sub-tile
// as 4 x 4 accumulators of v_mfma_f64_16x16x4f64.  A (128 x 16) and B
// (16 x 256) of each 16-deep K step are staged once in double-buffered LDS
// and shared by the eight waves; the next step's global loads are in flight
// during this step's MFMAs.  Per output and K step the workgroup moves 1.5
// bytes from global memory, against 4 for the 128 x 128 kernel that loads
// every wave's operands itself.  The A rows are stored permuted (LDS
// position 16 i + l of a wave's 64 rows holds its row 4 l + i), so MFMA block
// i's operand row l is C row 4 l + i and a lane's four outputs of element r
// are four consecutive rows: one 32-byte access of C each.  Requires m a
// multiple of 128, n of 256, k of 16, and 32-byte aligned columns.
constexpr int XM = 128, XN = 256, XK = 16, XT = 512;

template <int TRAILING>
__global__ __launch_bounds__(XT, 1) void dgemm_nn_sub_wide_kernel(
    int m, int n, int k, const double *__restrict__ A, int lda,
    const double *__restrict__ B, int ldb, double *__restrict__ C, int ldc,
    unsigned long long *stamp)
{
    __shared__ __align__(16) double As[2][XK][XM];
    __shared__ __align__(16) double Bs[2][XK][XN];
    launch_stamp_begin(stamp);
    const int tiles_m = m / XM, tiles_n = n / XN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    // global -> register: A rows a_r .. a_r + 3 at k = a_k; B column b_c,
    // k = b_k .. b_k + 7
    const int a_r = 4 * (tid & 31), a_k = tid >> 5;
    const int a_pos = (a_r / 64) * 64 + (a_r % 64) / 4;  // + 16 e for row a_r + e
    const int b_c = tid & 255, b_k = (tid >> 8) * 8;
    const int nk = k / XK;
    for (int tile = blockIdx.x; tile < tiles_m * tiles_n; tile += gridDim.x) {
        int tmi, tni;
        tile_coords(tile, tiles_m, tiles_n, tmi, tni);
        const int bm = tmi * XM, bn = tni * XN;
        const double *Ap = A + bm + a_r + (size_t)a_k * lda;
        const double *Bp = B + (size_t)(bn + b_c) * ldb + b_k;
        d4_t ra, rb[2];
        auto gload = [&](int kt) {
            const size_t k0 = (size_t)kt * XK;
            ra = *(const d4_t *)(Ap + k0 * lda);
            rb[0] = *(const d4_t *)(Bp + k0);
            rb[1] = *(const d4_t *)(Bp + k0 + 4);
        };
        auto lstore = [&](int buf) {
#pragma unroll
            for (int e = 0; e < 4; e++) As[buf][a_k][a_pos + 16 * e] = ra[e];
#pragma unroll
            for (int q = 0; q < 2; q++)
#pragma unroll
                for (int e = 0; e < 4; e++) Bs[buf][b_k + 4 * q + e][b_c] = rb[q][e];
        };
        d4_t acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = d4_t{0.0, 0.0, 0.0, 0.0};
        gload(0);
        lstore(0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt++) {
            const int buf = kt & 1;
            if (kt + 1 < nk) gload(kt + 1);
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) {
                double a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; i++) a[i] = As[buf][4 * s4 + lk][wm + 16 * i + lr];
#pragma unroll
                for (int j = 0; j < 4; j++) b[j] = Bs[buf][4 * s4 + lk][wn + 16 * j + lr];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            }
            if (kt + 1 < nk) lstore(buf ^ 1);
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double *Cc = C + (size_t)(bn + wn + 16 * j + lr) * ldc + bm + wm;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                d4_t *cp = (d4_t *)(Cc + 4 * (lk + 4 * r));
                d4_t c = *cp;
                c[0] -= acc[0][j][r];
                c[1] -= acc[1][j][r];
                c[2] -= acc[2][j][r];
                c[3] -= acc[3][j][r];
                *cp = c;
            }
        }
    }
    launch_stamp_end(stamp);
}

// C[m x n] -= A[m x k] * B[k x n] for n <= 16 (the right-hand sides of the
// triangular solves).  HBM-bound: each thread streams rows of A once.
template <int NMAX>
__global__ __launch_bounds__(256) void dgemm_nn_sub_skinny_kernel(
    int m, int n, int k, const double *__restrict__ A, int lda,
    const double *__restrict__ B, int ldb, double *__restrict__ C, int ldc)
{
    extern __shared__ double Bsh[];  // k x n, column-major with ld k
    for (int idx = threadIdx.x; idx < k * n; idx += blockDim.x) {
        const int kk = idx % k, jj = idx / k;
        Bsh[idx] = B[kk + (size_t)jj * ldb];
    }
    __syncthreads();
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m) return;
    double acc[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; j++) acc[j] = 0.0;
    const double *Ar = A + row;
    for (int kk = 0; kk < k; kk++) {
        const double a = Ar[(size_t)kk * lda];
#pragma unroll
        for (int j = 0; j < NMAX; j++)
            if (j < n) acc[j] = fma(a, Bsh[kk + j * k], acc[j]);
    }
#pragma unroll
    for (int j = 0; j < NMAX; j++)
        if (j < n) C[row + (size_t)j * ldc] -= acc[j];
}

}  // namespace

// The wide interior kernel serves the outer trailing update (SLA_DGEMM_WIDE=0
// in the environment selects the 128 x 128 kernel instead, for comparisons);
// the panel's GEMMs keep the 128 x 128 kernels, whose smaller footprint
// leaves room beside the concurrently running trailing update.
static bool dgemm_wide_on()
{
    static const bool on = [] {
        const char *e = getenv("SLA_DGEMM_WIDE");
        return !(e && e[0] == '0');
    }();
    return on;
}

void dgemm_nn_sub(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                  double *C, int ldc, hipStream_t s, bool trailing, int max_wgs,
                  unsigned long long *stamp)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (!trailing && gemm_small<double>(m, n, k, A, lda, B, ldb, C, ldc, s)) return;
    if (n <= 16 && (size_t)k * n * sizeof(double) <= 48 * 1024) {
        const int grid = ceil_div(m, 256);
        const size_t sh = (size_t)k * n * sizeof(double);
        if (n <= 1)
            hipLaunchKernelGGL(dgemm_nn_sub_skinny_kernel<1>, dim3(grid), dim3(256), sh, s, m, n, k, A, lda, B, ldb, C, ldc);
        else if (n <= 4)
            hipLaunchKernelGGL(dgemm_nn_sub_skinny_kernel<4>, dim3(grid), dim3(256), sh, s, m, n, k, A, lda, B, ldb, C, ldc);
        else
            hipLaunchKernelGGL(dgemm_nn_sub_skinny_kernel<16>, dim3(grid), dim3(256), sh, s, m, n, k, A, lda, B, ldb, C, ldc);
        SLA_LAUNCH_CHECK();
        return;
    }
    auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 31) == 0; };
    const bool aligned_all = k % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
                             aligned(A) && aligned(B) && aligned(C);
    const int mw = (m / XM) * XM, nw = (n / XN) * XN;
    if (trailing && dgemm_wide_on() && mw > 0 && nw > 0 && aligned_all) {
        const int tiles = (mw / XM) * (nw / XN);
        const int grid = (max_wgs >= 8 && tiles > max_wgs) ? (max_wgs / 8) * 8 : tiles;
        hipLaunchKernelGGL(dgemm_nn_sub_wide_kernel<1>, dim3(grid), dim3(XT), 0, s, mw, nw, k, A,
                           lda, B, ldb, C, ldc, stamp);
        SLA_LAUNCH_CHECK();
        // the rows below (all columns) and the columns right of the interior
        if (m > mw) dgemm_nn_sub(m - mw, n, k, A + mw, lda, B, ldb, C + mw, ldc, s, trailing,
                                 max_wgs, stamp);
        if (n > nw)
            dgemm_nn_sub(mw, n - nw, k, A, lda, B + (size_t)nw * ldb, ldb, C + (size_t)nw * ldc,
                         ldc, s, trailing, max_wgs, stamp);
        return;
    }
    const int mf = (m / 128) * 128, nf = (n / 128) * 128;
    if (mf > 0 && nf > 0 && k % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
        aligned(A) && aligned(B) && aligned(C)) {
        const int tiles = (mf / 128) * (nf / 128);
        // a persistent grid (multiple of 8) when the caller caps the workgroups
        const int grid = (max_wgs >= 8 && tiles > max_wgs) ? (max_wgs / 8) * 8 : tiles;
        if (trailing)
            hipLaunchKernelGGL(dgemm_nn_sub_direct_kernel<1>, dim3(grid), dim3(256), 0, s, mf, nf,
                               k, A, lda, B, ldb, C, ldc, stamp);
        else
            hipLaunchKernelGGL(dgemm_nn_sub_direct_kernel<0>, dim3(tiles), dim3(256), 0, s, mf, nf,
                               k, A, lda, B, ldb, C, ldc, (unsigned long long *)nullptr);
        SLA_LAUNCH_CHECK();
        // ragged edges: the rows below the interior (all columns) and the
        // columns right of it
        if (m > mf) dgemm_nn_sub(m - mf, n, k, A + mf, lda, B, ldb, C + mf, ldc, s, trailing,
                                 max_wgs);
        if (n > nf)
            dgemm_nn_sub(mf, n - nf, k, A, lda, B + (size_t)nf * ldb, ldb, C + (size_t)nf * ldc,
                         ldc, s, trailing, max_wgs);
        return;
    }
    const int tiles = ceil_div(m, TM) * ceil_div(n, TN);
    hipLaunchKernelGGL(dgemm_nn_sub_kernel, dim3(tiles), dim3(256), 0, s, m, n, k, A, lda, B, ldb,
                       C, ldc);
    SLA_LAUNCH_CHECK();
}

}  // namespace sla
