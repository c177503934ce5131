// gemm_lowp.hip -- the trailing update of the FP32-storage factorization used
// by the mixed-precision solver (a factorization in low precision, refined to
// FP64 accuracy): C[m x n] -= A[m x k] * B[k x n] with C in FP32 and the
// products in
//   * FP32 MFMA (v_mfma_f32_16x16x4_f32), or
//   * BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, operands
//     rounded to BF16 once per block column, or
//   * FP8-E4M3 MFMA (v_mfma_scale_f32_16x16x128_f8f6f4, unit block scales)
//     with FP32 accumulation, each operand block divided by its own scale
//     (max|block| / 448) before rounding and the product scaled back.
//   * split 16-bit products (BF16 or FP16 MFMA, v_mfma_f32_16x16x32_{bf16,f16}):
//     every FP32 operand is a sum of S 16-bit slices and the product keeps
//     the slice products of order below S (3 MFMAs for S = 2, 6 for S = 3),
//     which restores most of the FP32 product's accuracy at the 16-bit rate.
// It is the same dgemm_('N','N', ..., -1, L21, U12, +1, A22) of PDGETRF
// (SRC/pdgetrf.f:288-290), in lower precision.
//
// Lane layouts (checked on gfx950 with one-hot probes):
//   f32 16x16x4 : A lane l = A[l%16][l/16], B = B[l/16][l%16],
//                 D element r of lane l = D[4*(l/16) + r][l%16]
//   bf16 16x16x32: A lane l element e = A[l%16][8*(l/16) + e],
//                 B = B[8*(l/16) + e][l%16], D as f32 16x16x4.
//   f8 16x16x128 : A lane l byte e = A[l%16][32*(l/16) + e],
//                 B = B[32*(l/16) + e][l%16], D as f32 16x16x4.
#include "sla_common.h"
#include "sla_kernels.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace sla {

namespace {

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef short s8_t __attribute__((ext_vector_type(8)));
typedef int i8_t __attribute__((ext_vector_type(8)));

// E4M3 (OCP) has a largest finite value of 448; a block is divided by
// max|block| / 448 before rounding so that its range fills the format.
constexpr float E4M3_MAX = 448.0f;
__device__ __forceinline__ float fp8_scale(float absmax)
{
    return absmax > 0.f ? absmax / E4M3_MAX : 1.0f;
}

template <typename TO>
__device__ __forceinline__ TO to_lowp(float v);
template <>
__device__ __forceinline__ __bf16 to_lowp<__bf16>(float v)
{
    return (__bf16)v;  // round to nearest even
}
template <>
__device__ __forceinline__ uint8_t to_lowp<uint8_t>(float v)
{
    // saturating round-to-nearest conversion of the hardware (E4M3 on gfx950)
    v = fminf(fmaxf(v, -E4M3_MAX), E4M3_MAX);
    return (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v, 0.f, 0, false) & 0xff);
}

__device__ __forceinline__ void tile_coords_lp(int id, int tiles_m, int tiles_n, int &tm, int &tn)
{
    const int total = tiles_m * tiles_n;
    int lid = id;
    if (total >= 64) {
        const int per = total / 8, rem = total % 8;
        const int xcd = id % 8, k = id / 8;
        const int start = xcd * per + (xcd < rem ? xcd : rem);
        const int len = per + (xcd < rem ? 1 : 0);
        lid = (k < len) ? start + k : id;
    }
    const int band = 8, bands_full = tiles_n / band, in_full = bands_full * band * tiles_m;
    if (lid < in_full) {
        const int b = lid / (band * tiles_m), r = lid % (band * tiles_m);
        tm = r % tiles_m;
        tn = b * band + r / tiles_m;
    } else {
        const int r = lid - in_full;
        tm = r % tiles_m;
        tn = bands_full * band + r / tiles_m;
    }
}

// ---------------------------------------------------------------------------
// FP32, LDS-free, 128 x 128 per 256-thread workgroup, 64 x 64 per wave.
// Row permutation: block i, operand row lr <-> C row wm + 4*lr + i;
// K permutation: k-lane lk, step s <-> k = k0 + 4*lk + s.  A lane therefore
// loads 4 consecutive rows (a float4) per k and 4 consecutive k (a float4)
// per column.  D element r (block row 4*lk + r) is C row wm + 16*lk + 4*r + i.
constexpr int SDEPTH = 3;

// TRAILING: 1 for the outer trailing update's launches (kernel traces tell
// them apart), 0 otherwise; the code is the same.
template <int TRAILING>
__global__ __launch_bounds__(256, 2) void sgemm_nn_sub_direct_kernel(
    int m, int n, int k, const float *__restrict__ A, int lda, const float *__restrict__ B,
    int ldb, float *__restrict__ C, int ldc, unsigned long long *stamp)
{
    launch_stamp_begin(stamp);
    // persistent when the grid is smaller than the tile count (see
    // dgemm_nn_sub_direct_kernel)
    for (int tile = blockIdx.x; tile < (m / 128) * (n / 128); tile += gridDim.x) {
    int tmi, tni;
    tile_coords_lp(tile, m / 128, n / 128, tmi, tni);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = tmi * 128 + (wave & 1) * 64, wn = tni * 128 + (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    const float *Ap = A + (wm + 4 * lr) + (size_t)(4 * lk) * lda;
    const float *Bp = B + (size_t)(wn + lr) * ldb + 4 * lk;
    const size_t bstep = (size_t)16 * ldb;
    f4_t a[SDEPTH][4], b[SDEPTH][4];
    auto load = [&](int t, int q) {
        const size_t k0 = (size_t)16 * t;
#pragma unroll
        for (int s = 0; s < 4; s++) a[q][s] = *(const f4_t *)(Ap + (k0 + s) * lda);
#pragma unroll
        for (int j = 0; j < 4; j++) b[q][j] = *(const f4_t *)(Bp + j * bstep + k0);
    };
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    const int nk = k / 16;
#pragma unroll
    for (int q = 0; q < SDEPTH - 1; q++)
        if (q < nk) load(q, q);
    for (int t = 0; t < nk; t += SDEPTH) {
#pragma unroll
        for (int q = 0; q < SDEPTH; q++) {
            if (t + q < nk) {
                const int nt = t + q + SDEPTH - 1;
                if (nt < nk) load(nt, (q + SDEPTH - 1) % SDEPTH);
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q][s][i], b[q][j][s],
                                                                            acc[i][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float *Cc = C + (size_t)(wn + 16 * j + lr) * ldc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            f4_t *cp = (f4_t *)(Cc + wm + 16 * lk + 4 * r);
            f4_t c = *cp;
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

// Ragged edges (and small shapes) of either precision: one thread per element
// of C, operands read through a converter.  Only a thin border ever lands here.
template <typename TA>
__device__ __forceinline__ float to_f(TA v) { return (float)v; }

template <typename TA, bool AROW>
__global__ __launch_bounds__(256) void gemm_lp_edge_kernel(int m, int n, int k, const TA *A,
                                                           int lda, const TA *B, int ldb,
                                                           float *C, int ldc)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= m || j >= n) return;
    float acc = 0.f;
    for (int kk = 0; kk < k; kk++) {
        const float av = AROW ? to_f(A[(size_t)i * lda + kk]) : to_f(A[i + (size_t)kk * lda]);
        acc = fmaf(av, to_f(B[kk + (size_t)j * ldb]), acc);
    }
    C[i + (size_t)j * ldc] -= acc;
}

// ---------------------------------------------------------------------------
// BF16 operands packed once per block column:
//   Ap : the m x k multiplier block L21, row-major (k contiguous, ld = ldak)
//   Bp : the k x n block row U12, column-major (k contiguous, ld = ldbk)
// so each lane's 8 k-values are one 16-byte load.  128 x 128 per workgroup,
// 64 x 64 per wave: 4 x 4 accumulators of 16 x 16.
constexpr int BDEPTH = 3;

__global__ __launch_bounds__(256, 2) void gemm_bf16_nn_sub_direct_kernel(
    int m, int n, int k, const __bf16 *__restrict__ Ap, int ldak, const __bf16 *__restrict__ Bp,
    int ldbk, float *__restrict__ C, int ldc, unsigned long long *stamp)
{
    launch_stamp_begin(stamp);
    int tmi, tni;
    tile_coords_lp(blockIdx.x, m / 128, n / 128, tmi, tni);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = tmi * 128 + (wave & 1) * 64, wn = tni * 128 + (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    const __bf16 *Ar = Ap + (size_t)(wm + lr) * ldak + 8 * lk;
    const __bf16 *Bc = Bp + (size_t)(wn + lr) * ldbk + 8 * lk;
    const size_t astep = (size_t)16 * ldak, bstep = (size_t)16 * ldbk;
    s8_t a[BDEPTH][4], b[BDEPTH][4];
    auto load = [&](int t, int q) {
        const size_t k0 = (size_t)32 * t;
#pragma unroll
        for (int i = 0; i < 4; i++) a[q][i] = *(const s8_t *)(Ar + i * astep + k0);
#pragma unroll
        for (int j = 0; j < 4; j++) b[q][j] = *(const s8_t *)(Bc + j * bstep + k0);
    };
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    const int nk = k / 32;
#pragma unroll
    for (int q = 0; q < BDEPTH - 1; q++)
        if (q < nk) load(q, q);
    for (int t = 0; t < nk; t += BDEPTH) {
#pragma unroll
        for (int q = 0; q < BDEPTH; q++) {
            if (t + q < nk) {
                const int nt = t + q + BDEPTH - 1;
                if (nt < nk) load(nt, (q + BDEPTH - 1) % BDEPTH);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8_t, a[q][i]), __builtin_bit_cast(bf8_t, b[q][j]),
                            acc[i][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float *Cc = C + (size_t)(wn + 16 * j + lr) * ldc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            f4_t *cp = (f4_t *)(Cc + wm + 16 * i + 4 * lk);
            f4_t c = *cp;
            c[0] -= acc[i][j][0];
            c[1] -= acc[i][j][1];
            c[2] -= acc[i][j][2];
            c[3] -= acc[i][j][3];
            *cp = c;
        }
    }
    launch_stamp_end(stamp);
}

// ---------------------------------------------------------------------------
// FP8-E4M3 operands, packed like the BF16 ones (k contiguous, ld multiple of
// 128), on v_mfma_scale_f32_16x16x128_f8f6f4 with unit block scales: lane l
// holds 32 consecutive k values of row (column) l%16 starting at 32*(l/16);
// D as f32 16x16x4.  The products are multiplied back by sa * sb.
constexpr int FDEPTH = 2;

__global__ __launch_bounds__(256, 2) void gemm_fp8_nn_sub_direct_kernel(
    int m, int n, int k, const uint8_t *__restrict__ Ap, int ldak, const uint8_t *__restrict__ Bp,
    int ldbk, const float *__restrict__ scales, float *__restrict__ C, int ldc,
    unsigned long long *stamp)
{
    launch_stamp_begin(stamp);
    int tmi, tni;
    tile_coords_lp(blockIdx.x, m / 128, n / 128, tmi, tni);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = tmi * 128 + (wave & 1) * 64, wn = tni * 128 + (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    const uint8_t *Ar = Ap + (size_t)(wm + lr) * ldak + 32 * lk;
    const uint8_t *Bc = Bp + (size_t)(wn + lr) * ldbk + 32 * lk;
    const size_t astep = (size_t)16 * ldak, bstep = (size_t)16 * ldbk;
    i8_t a[FDEPTH][4], b[FDEPTH][4];
    auto load = [&](int t, int q) {
        const size_t k0 = (size_t)128 * t;
#pragma unroll
        for (int i = 0; i < 4; i++) a[q][i] = *(const i8_t *)(Ar + i * astep + k0);
#pragma unroll
        for (int j = 0; j < 4; j++) b[q][j] = *(const i8_t *)(Bc + j * bstep + k0);
    };
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    const int nk = k / 128;
#pragma unroll
    for (int q = 0; q < FDEPTH - 1; q++)
        if (q < nk) load(q, q);
    for (int t = 0; t < nk; t += FDEPTH) {
#pragma unroll
        for (int q = 0; q < FDEPTH; q++) {
            if (t + q < nk) {
                const int nt = t + q + FDEPTH - 1;
                if (nt < nk) load(nt, (q + FDEPTH - 1) % FDEPTH);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                            a[q][i], b[q][j], acc[i][j], 0, 0, 0, 0, 0, 0);
            }
        }
    }
    const float sab = fp8_scale(scales[0]) * fp8_scale(scales[1]);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float *Cc = C + (size_t)(wn + 16 * j + lr) * ldc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            f4_t *cp = (f4_t *)(Cc + wm + 16 * i + 4 * lk);
            f4_t c = *cp;
            c[0] -= sab * acc[i][j][0];
            c[1] -= sab * acc[i][j][1];
            c[2] -= sab * acc[i][j][2];
            c[3] -= sab * acc[i][j][3];
            *cp = c;
        }
    }
    launch_stamp_end(stamp);
}

// ---------------------------------------------------------------------------
// Split 16-bit products.  Each FP32 operand x (FP16: first scaled by 2^e, e
// chosen per operand block so that max|x| 2^e lies in [2^14, 2^15)) is written
// as x = x0 + x1 (+ x2) with x0 = r16(x), x1 = r16(x - x0), x2 = r16(x - x0 -
// x1) (each difference is exact in FP32), r16 = round to nearest even in the
// 16-bit format.  The product sums the slice products a_i b_j with i + j < S:
//   S = 2: a0 b0 + a0 b1 + a1 b0; the dropped a1 b1 and the representation
//          error are ~2^-16 (BF16) or ~2^-21 (FP16) of |a| |b|;
//   S = 3: adds a0 b2 + a1 b1 + a2 b0; BF16 error ~2^-24 of |a| |b|, the FP32
//          product's own level.
// Packed operands: S planes, plane s of A row-major (k contiguous, ld ldak),
// plane s of B column-major (k contiguous, ld ldbk), planes aplane / bplane
// elements apart.  The smaller slice products go first into the FP32
// accumulator.
template <typename T16>
__device__ __forceinline__ f4_t mfma16(s8_t a, s8_t b, f4_t c);
template <>
__device__ __forceinline__ f4_t mfma16<__bf16>(s8_t a, s8_t b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8_t, a),
                                                   __builtin_bit_cast(bf8_t, b), c, 0, 0, 0);
}
template <>
__device__ __forceinline__ f4_t mfma16<_Float16>(s8_t a, s8_t b, f4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8_t, a),
                                                  __builtin_bit_cast(h8_t, b), c, 0, 0, 0);
}

// Power-of-two exponent of an FP16 operand block with max|x| = absmax:
// absmax 2^e lies in [2^14, 2^15) (0 for a zero block).  BF16 has FP32's
// range and is not scaled.
template <typename T16>
__device__ __forceinline__ int split_exp(float absmax)
{
    if (std::is_same<T16, __bf16>::value || !(absmax > 0.f)) return 0;
    int e = 0;
    (void)frexpf(absmax, &e);  // absmax = f 2^e, f in [0.5, 1)
    return 15 - e;
}

template <typename T16, int S>
__device__ __forceinline__ void split16(float x, T16 (&out)[S])
{
#pragma unroll
    for (int i = 0; i < S; i++) {
        out[i] = (T16)x;
        x -= (float)out[i];
    }
}

constexpr int SPDEPTH = 2;

template <typename T16, int S>
__global__ __launch_bounds__(256, S == 2 ? 2 : 1) void gemm_split_nn_sub_direct_kernel(
    int m, int n, int k, const T16 *__restrict__ Ap, int ldak, size_t aplane,
    const T16 *__restrict__ Bp, int ldbk, size_t bplane, const float *__restrict__ absmax,
    float *__restrict__ C, int ldc, unsigned long long *stamp)
{
    launch_stamp_begin(stamp);
    int tmi, tni;
    tile_coords_lp(blockIdx.x, m / 128, n / 128, tmi, tni);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = tmi * 128 + (wave & 1) * 64, wn = tni * 128 + (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    const T16 *Ar = Ap + (size_t)(wm + lr) * ldak + 8 * lk;
    const T16 *Bc = Bp + (size_t)(wn + lr) * ldbk + 8 * lk;
    const size_t astep = (size_t)16 * ldak, bstep = (size_t)16 * ldbk;
    s8_t a[SPDEPTH][S][4], b[SPDEPTH][S][4];
    auto load = [&](int t, int q) {
        const size_t k0 = (size_t)32 * t;
#pragma unroll
        for (int sl = 0; sl < S; sl++) {
#pragma unroll
            for (int i = 0; i < 4; i++)
                a[q][sl][i] = *(const s8_t *)(Ar + sl * aplane + i * astep + k0);
#pragma unroll
            for (int j = 0; j < 4; j++)
                b[q][sl][j] = *(const s8_t *)(Bc + sl * bplane + j * bstep + k0);
        }
    };
    f4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};
    const int nk = k / 32;
#pragma unroll
    for (int q = 0; q < SPDEPTH - 1; q++)
        if (q < nk) load(q, q);
    for (int t = 0; t < nk; t += SPDEPTH) {
#pragma unroll
        for (int q = 0; q < SPDEPTH; q++) {
            if (t + q < nk) {
                const int nt = t + q + SPDEPTH - 1;
                if (nt < nk) load(nt, (q + SPDEPTH - 1) % SPDEPTH);
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        f4_t c = acc[i][j];
                        if constexpr (S == 3) {
                            c = mfma16<T16>(a[q][2][i], b[q][0][j], c);
                            c = mfma16<T16>(a[q][1][i], b[q][1][j], c);
                            c = mfma16<T16>(a[q][0][i], b[q][2][j], c);
                        }
                        c = mfma16<T16>(a[q][1][i], b[q][0][j], c);
                        c = mfma16<T16>(a[q][0][i], b[q][1][j], c);
                        acc[i][j] = mfma16<T16>(a[q][0][i], b[q][0][j], c);
                    }
            }
        }
    }
    const float back = ldexpf(1.0f, -(split_exp<T16>(absmax[0]) + split_exp<T16>(absmax[1])));
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float *Cc = C + (size_t)(wn + 16 * j + lr) * ldc;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            f4_t *cp = (f4_t *)(Cc + wm + 16 * i + 4 * lk);
            f4_t c = *cp;
            c[0] -= back * acc[i][j][0];
            c[1] -= back * acc[i][j][1];
            c[2] -= back * acc[i][j][2];
            c[3] -= back * acc[i][j][3];
            *cp = c;
        }
    }
    launch_stamp_end(stamp);
}

// The split product for large tiles: a TM x 128 tile of C per 256-thread
// workgroup (TM = 256 for two slices, 128 for three), the operand slices of
// each 32-deep K step staged once in LDS and shared by the four waves (wave
// w: rows (TM / 2) (w & 1) .., columns 64 (w >> 1) ..: TM / 32 x 4
// accumulators of 16 x 16), the next K step's slices loaded into registers
// while the MFMAs of this one run, LDS double buffered (one workgroup
// barrier per K step).  LDS rows hold 32 k values plus 8 of padding (80
// bytes), so the 16 lanes that read one 16-byte fragment each cover distinct
// banks.  Either slice count takes 120 KB of LDS; the products run in the
// direct kernel's order, so both kernels give the same bits.
constexpr int SL_N = 128, SL_K = 32, SL_LD = SL_K + 8;
constexpr int split_tile_m(int S) { return S == 2 ? 256 : 128; }
constexpr size_t split_lds_bytes(int S)
{
    return (size_t)2 * S * (split_tile_m(S) + SL_N) * SL_LD * 2;
}

template <typename T16, int S>
__global__ __launch_bounds__(256, 1) void gemm_split_lds_kernel(
    int m, int n, int k, const T16 *__restrict__ Ap, int ldak, size_t aplane,
    const T16 *__restrict__ Bp, int ldbk, size_t bplane, const float *__restrict__ absmax,
    float *__restrict__ C, int ldc, unsigned long long *stamp)
{
    constexpr int TM = split_tile_m(S), MB = TM / 32, AR = TM / 64;
    extern __shared__ __align__(16) unsigned char sl_lds[];
    // [buf][slice][row][SL_LD] of A, then [buf][slice][col][SL_LD] of B
    T16 *As = reinterpret_cast<T16 *>(sl_lds);
    T16 *Bs = As + (size_t)2 * S * TM * SL_LD;
    auto a_at = [&](int buf, int sl, int row) { return As + (((size_t)buf * S + sl) * TM + row) * SL_LD; };
    auto b_at = [&](int buf, int sl, int col) { return Bs + (((size_t)buf * S + sl) * SL_N + col) * SL_LD; };

    launch_stamp_begin(stamp);
    const int tiles_m = m / TM, tiles_n = n / SL_N;
    int tmi, tni;
    tile_coords_lp(blockIdx.x, tiles_m, tiles_n, tmi, tni);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = tmi * TM, col0 = tni * SL_N;
    const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;

    // global -> register staging: A rows tid / 4 + 64 i (i < TM / 64), B
    // columns tid / 4 + 64 i (i < 2), 16-byte chunk tid % 4 of the 32 k values
    const int srow = tid >> 2, sch = (tid & 3) * 8;
    s8_t ra[S][AR], rb[S][2];
    auto gload = [&](int t) {
        const size_t k0 = (size_t)SL_K * t + sch;
#pragma unroll
        for (int sl = 0; sl < S; sl++) {
#pragma unroll
            for (int i = 0; i < AR; i++)
                ra[sl][i] = *(const s8_t *)(Ap + sl * aplane + (size_t)(row0 + srow + 64 * i) * ldak + k0);
#pragma unroll
            for (int i = 0; i < 2; i++)
                rb[sl][i] = *(const s8_t *)(Bp + sl * bplane + (size_t)(col0 + srow + 64 * i) * ldbk + k0);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int sl = 0; sl < S; sl++) {
#pragma unroll
            for (int i = 0; i < AR; i++) *(s8_t *)(a_at(buf, sl, srow + 64 * i) + sch) = ra[sl][i];
#pragma unroll
            for (int i = 0; i < 2; i++) *(s8_t *)(b_at(buf, sl, srow + 64 * i) + sch) = rb[sl][i];
        }
    };

    f4_t acc[MB][4];
#pragma unroll
    for (int i = 0; i < MB; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4_t{0.f, 0.f, 0.f, 0.f};

    const int nk = k / SL_K;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        const int buf = t & 1;
        if (t + 1 < nk) gload(t + 1);
        s8_t fb[S][4];
#pragma unroll
        for (int sl = 0; sl < S; sl++)
#pragma unroll
            for (int j = 0; j < 4; j++) fb[sl][j] = *(const s8_t *)(b_at(buf, sl, wn + 16 * j + lr) + 8 * lk);
        s8_t fa[MB][S];
#pragma unroll
        for (int i = 0; i < MB; i++)
#pragma unroll
            for (int sl = 0; sl < S; sl++)
                fa[i][sl] = *(const s8_t *)(a_at(buf, sl, wm + 16 * i + lr) + 8 * lk);
        // product by product over all accumulators: every accumulator gets
        // its products in the direct kernel's order, and consecutive MFMAs
        // never wait for each other's result
        auto prod = [&](int sa, int sb) {
#pragma unroll
            for (int i = 0; i < MB; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = mfma16<T16>(fa[i][sa], fb[sb][j], acc[i][j]);
        };
        if constexpr (S == 3) {
            prod(2, 0);
            prod(1, 1);
            prod(0, 2);
        }
        prod(1, 0);
        prod(0, 1);
        prod(0, 0);
        if (t + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }

    const float back = ldexpf(1.0f, -(split_exp<T16>(absmax[0]) + split_exp<T16>(absmax[1])));
#pragma unroll
    for (int j = 0; j < 4; j++) {
        float *Cc = C + (size_t)(col0 + wn + 16 * j + lr) * ldc + row0 + wm;
#pragma unroll
        for (int i = 0; i < MB; i++) {
            f4_t *cp = (f4_t *)(Cc + 16 * i + 4 * lk);
            f4_t c = *cp;
            c[0] -= back * acc[i][j][0];
            c[1] -= back * acc[i][j][1];
            c[2] -= back * acc[i][j][2];
            c[3] -= back * acc[i][j][3];
            *cp = c;
        }
    }
    launch_stamp_end(stamp);
}

// Ragged edges of the split form: one thread per element of C, the same
// slice products in the same order per k.
template <typename T16, int S>
__global__ __launch_bounds__(256) void gemm_split_edge_kernel(
    int m, int n, int k, const T16 *A, int lda, size_t aplane, const T16 *B, int ldb,
    size_t bplane, const float *absmax, float *C, int ldc)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= m || j >= n) return;
    float acc = 0.f;
    for (int kk = 0; kk < k; kk++) {
        float av[S], bv[S];
#pragma unroll
        for (int sl = 0; sl < S; sl++) {
            av[sl] = (float)A[sl * aplane + (size_t)i * lda + kk];
            bv[sl] = (float)B[sl * bplane + kk + (size_t)j * ldb];
        }
        if constexpr (S == 3) {
            acc = fmaf(av[2], bv[0], acc);
            acc = fmaf(av[1], bv[1], acc);
            acc = fmaf(av[0], bv[2], acc);
        }
        acc = fmaf(av[1], bv[0], acc);
        acc = fmaf(av[0], bv[1], acc);
        acc = fmaf(av[0], bv[0], acc);
    }
    const float back = ldexpf(1.0f, -(split_exp<T16>(absmax[0]) + split_exp<T16>(absmax[1])));
    C[i + (size_t)j * ldc] -= back * acc;
}

// L21 (m x k, column-major) -> S row-major planes (ld kpad), zero padded.
template <typename T16, int S>
__global__ void pack_rows_split_kernel(int m, int k, int kpad, const float *__restrict__ A, int lda,
                                       const float *__restrict__ absmax, T16 *__restrict__ P,
                                       size_t plane)
{
    __shared__ float t[32][33];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int e = split_exp<T16>(absmax ? *absmax : 0.f);
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + tx, kk = k0 + y;
        t[y][tx] = (r < m && kk < k) ? ldexpf(A[r + (size_t)kk * lda], e) : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + y, kk = k0 + tx;
        if (r < m && kk < kpad) {
            T16 sl[S];
            split16<T16, S>(t[tx][y], sl);
#pragma unroll
            for (int q = 0; q < S; q++) P[q * plane + (size_t)r * kpad + kk] = sl[q];
        }
    }
}

// U12 (k x n, column-major) -> S column-major planes (ld kpad), zero padded.
template <typename T16, int S>
__global__ void pack_cols_split_kernel(int k, int kpad, int n, const float *__restrict__ B, int ldb,
                                       const float *__restrict__ absmax, T16 *__restrict__ P,
                                       size_t plane)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)kpad * n) return;
    const int e = split_exp<T16>(absmax ? *absmax : 0.f);
    const int kk = (int)(idx % kpad), j = (int)(idx / kpad);
    T16 sl[S];
    split16<T16, S>(kk < k ? ldexpf(B[kk + (size_t)j * ldb], e) : 0.f, sl);
#pragma unroll
    for (int q = 0; q < S; q++) P[q * plane + idx] = sl[q];
}

__device__ __forceinline__ float fp8_value(uint8_t b)
{
    const int e = (b >> 3) & 15, mnt = b & 7;
    const float v = e == 0 ? ldexpf((float)mnt * 0.125f, -6) : ldexpf(1.0f + mnt * 0.125f, e - 7);
    return (b & 0x80) ? -v : v;
}

// Ragged edges of the FP8 form (row-major A, column-major B, both k-contiguous).
__global__ __launch_bounds__(256) void gemm_fp8_edge_kernel(int m, int n, int k,
                                                            const uint8_t *A, int lda,
                                                            const uint8_t *B, int ldb,
                                                            const float *scales, float *C, int ldc)
{
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= m || j >= n) return;
    float acc = 0.f;
    for (int kk = 0; kk < k; kk++)
        acc = fmaf(fp8_value(A[(size_t)i * lda + kk]), fp8_value(B[kk + (size_t)j * ldb]), acc);
    C[i + (size_t)j * ldc] -= fp8_scale(scales[0]) * fp8_scale(scales[1]) * acc;
}

// ---- packing and conversions ------------------------------------------------
// L21 (m x k, column-major, ld lda, FP32) -> row-major (ld kpad), zero padded.
template <typename TO, bool SCALED>
__global__ void pack_rows_kernel(int m, int k, int kpad, const float *__restrict__ A, int lda,
                                 const float *__restrict__ scale, TO *__restrict__ P)
{
    __shared__ float t[32][33];
    const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const float inv = SCALED ? 1.0f / fp8_scale(*scale) : 1.0f;
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + tx, kk = k0 + y;
        t[y][tx] = (r < m && kk < k) ? A[r + (size_t)kk * lda] * inv : 0.f;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
        const int r = r0 + y, kk = k0 + tx;
        if (r < m && kk < kpad) P[(size_t)r * kpad + kk] = to_lowp<TO>(t[tx][y]);
    }
}

// U12 (k x n, column-major, ld ldb, FP32) -> column-major (ld kpad), zero padded.
template <typename TO, bool SCALED>
__global__ void pack_cols_kernel(int k, int kpad, int n, const float *__restrict__ B, int ldb,
                                 const float *__restrict__ scale, TO *__restrict__ P)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)kpad * n) return;
    const float inv = SCALED ? 1.0f / fp8_scale(*scale) : 1.0f;
    const int kk = (int)(idx % kpad), j = (int)(idx / kpad);
    P[idx] = to_lowp<TO>(kk < k ? B[kk + (size_t)j * ldb] * inv : 0.f);
}

// Bit pattern of max |A| over an m x n block into *scale (atomic max on the
// bits, which order like the values for non-negative floats).
__global__ __launch_bounds__(256) void absmax_kernel(int m, int n, const float *__restrict__ A,
                                                     int lda, unsigned *__restrict__ out)
{
    float v = 0.f;
    const size_t total = (size_t)m * n, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int i = (int)(idx % m), j = (int)(idx / m);
        v = fmaxf(v, fabsf(A[i + (size_t)j * lda]));
    }
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(v));
}

__global__ void convert_f64_f32_kernel(int m, int n, const double *__restrict__ A, int lda,
                                       float *__restrict__ B, int ldb, int *overflow)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    const double a = A[i + (size_t)j * lda];
    if (!(fabs(a) <= 3.4028234663852886e38)) *overflow = 1;  // also NaN
    B[i + (size_t)j * ldb] = (float)a;
}

}  // namespace

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

void sgemm_nn_sub(int m, int n, int k, const float *A, int lda, const float *B, int ldb, float *C,
                  int ldc, hipStream_t s, bool trailing, int max_wgs, int reserve_lds,
                  unsigned long long *stamp)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (!trailing && gemm_small<float>(m, n, k, A, lda, B, ldb, C, ldc, s)) return;
    const int mf = (m / 128) * 128, nf = (n / 128) * 128;
    if (mf > 0 && nf > 0 && k % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
        aligned16(A) && aligned16(B) && aligned16(C)) {
        const int tiles = (mf / 128) * (nf / 128);
        const int grid = (max_wgs >= 8 && tiles > max_wgs) ? (max_wgs / 8) * 8 : tiles;
        if (trailing) {
            // reserve_lds: LDS the kernel does not use, reserved so that at
            // most floor(LDS per CU / reserve_lds) workgroups share a CU and
            // registers stay free for the panel running beside the update
            if (reserve_lds > 0) {
                static int attr_set = 0;
                if (reserve_lds > attr_set) {
                    SLA_HIP(hipFuncSetAttribute(
                        reinterpret_cast<const void *>(&sgemm_nn_sub_direct_kernel<1>),
                        hipFuncAttributeMaxDynamicSharedMemorySize, reserve_lds));
                    attr_set = reserve_lds;
                }
            }
            hipLaunchKernelGGL(sgemm_nn_sub_direct_kernel<1>, dim3(grid), dim3(256),
                               (unsigned)std::max(0, reserve_lds), s, mf, nf, k, A, lda, B, ldb, C,
                               ldc, stamp);
        }
        else
            hipLaunchKernelGGL(sgemm_nn_sub_direct_kernel<0>, dim3((mf / 128) * (nf / 128)),
                               dim3(256), 0, s, mf, nf, k, A, lda, B, ldb, C, ldc,
                               (unsigned long long *)nullptr);
        SLA_LAUNCH_CHECK();
        if (m > mf)
            sgemm_nn_sub(m - mf, n, k, A + mf, lda, B, ldb, C + mf, ldc, s, trailing, max_wgs,
                         reserve_lds);
        if (n > nf)
            sgemm_nn_sub(mf, n - nf, k, A, lda, B + (size_t)nf * ldb, ldb, C + (size_t)nf * ldc,
                         ldc, s, trailing, max_wgs, reserve_lds);
        return;
    }
    hipLaunchKernelGGL((gemm_lp_edge_kernel<float, false>), dim3(ceil_div(m, 16), ceil_div(n, 16)),
                       dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
    SLA_LAUNCH_CHECK();
}

void gemm_bf16_nn_sub(int m, int n, int k, const __bf16 *Ap, int ldak, const __bf16 *Bp, int ldbk,
                      float *C, int ldc, hipStream_t s, unsigned long long *stamp)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int mf = (m / 128) * 128, nf = (n / 128) * 128;
    if (mf > 0 && nf > 0 && k % 32 == 0 && ldak % 8 == 0 && ldbk % 8 == 0 && ldc % 4 == 0 &&
        aligned16(Ap) && aligned16(Bp) && aligned16(C)) {
        hipLaunchKernelGGL(gemm_bf16_nn_sub_direct_kernel, dim3((mf / 128) * (nf / 128)), dim3(256),
                           0, s, mf, nf, k, Ap, ldak, Bp, ldbk, C, ldc, stamp);
        SLA_LAUNCH_CHECK();
        if (m > mf)
            gemm_bf16_nn_sub(m - mf, n, k, Ap + (size_t)mf * ldak, ldak, Bp, ldbk, C + mf, ldc, s);
        if (n > nf)
            gemm_bf16_nn_sub(mf, n - nf, k, Ap, ldak, Bp + (size_t)nf * ldbk, ldbk,
                             C + (size_t)nf * ldc, ldc, s);
        return;
    }
    hipLaunchKernelGGL((gemm_lp_edge_kernel<__bf16, true>), dim3(ceil_div(m, 16), ceil_div(n, 16)),
                       dim3(256), 0, s, m, n, k, Ap, ldak, Bp, ldbk, C, ldc);
    SLA_LAUNCH_CHECK();
}

template <typename TO, bool SCALED>
static void pack_rows(int m, int k, int kpad, const float *A, int lda, const float *scale, TO *P,
                      hipStream_t s)
{
    if (m <= 0 || kpad <= 0) return;
    hipLaunchKernelGGL((pack_rows_kernel<TO, SCALED>), dim3(ceil_div(m, 32), ceil_div(kpad, 32)),
                       dim3(256), 0, s, m, k, kpad, A, lda, scale, P);
    SLA_LAUNCH_CHECK();
}

template <typename TO, bool SCALED>
static void pack_cols(int k, int kpad, int n, const float *B, int ldb, const float *scale, TO *P,
                      hipStream_t s)
{
    if (n <= 0 || kpad <= 0) return;
    const size_t total = (size_t)kpad * n;
    hipLaunchKernelGGL((pack_cols_kernel<TO, SCALED>), dim3((unsigned)((total + 255) / 256)),
                       dim3(256), 0, s, k, kpad, n, B, ldb, scale, P);
    SLA_LAUNCH_CHECK();
}

void pack_rows_bf16(int m, int k, int kpad, const float *A, int lda, __bf16 *P, hipStream_t s)
{
    pack_rows<__bf16, false>(m, k, kpad, A, lda, nullptr, P, s);
}

void pack_cols_bf16(int k, int kpad, int n, const float *B, int ldb, __bf16 *P, hipStream_t s)
{
    pack_cols<__bf16, false>(k, kpad, n, B, ldb, nullptr, P, s);
}

void pack_rows_fp8(int m, int k, int kpad, const float *A, int lda, const float *scale,
                   uint8_t *P, hipStream_t s)
{
    pack_rows<uint8_t, true>(m, k, kpad, A, lda, scale, P, s);
}

void pack_cols_fp8(int k, int kpad, int n, const float *B, int ldb, const float *scale,
                   uint8_t *P, hipStream_t s)
{
    pack_cols<uint8_t, true>(k, kpad, n, B, ldb, scale, P, s);
}

void absmax_scale(int m, int n, const float *A, int lda, float *scale, hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    const size_t total = (size_t)m * n;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(256), 0, s, m, n, A, lda,
                       reinterpret_cast<unsigned *>(scale));
    SLA_LAUNCH_CHECK();
}

void gemm_fp8_nn_sub(int m, int n, int k, const uint8_t *Ap, int ldak, const uint8_t *Bp,
                     int ldbk, const float *scales, float *C, int ldc, hipStream_t s,
                     unsigned long long *stamp)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    const int mf = (m / 128) * 128, nf = (n / 128) * 128;
    if (mf > 0 && nf > 0 && k % 128 == 0 && ldak % 32 == 0 && ldbk % 32 == 0 && ldc % 4 == 0 &&
        aligned16(Ap) && aligned16(Bp) && aligned16(C)) {
        hipLaunchKernelGGL(gemm_fp8_nn_sub_direct_kernel, dim3((mf / 128) * (nf / 128)), dim3(256),
                           0, s, mf, nf, k, Ap, ldak, Bp, ldbk, scales, C, ldc, stamp);
        SLA_LAUNCH_CHECK();
        if (m > mf)
            gemm_fp8_nn_sub(m - mf, n, k, Ap + (size_t)mf * ldak, ldak, Bp, ldbk, scales, C + mf,
                            ldc, s);
        if (n > nf)
            gemm_fp8_nn_sub(mf, n - nf, k, Ap, ldak, Bp + (size_t)nf * ldbk, ldbk, scales,
                            C + (size_t)nf * ldc, ldc, s);
        return;
    }
    hipLaunchKernelGGL(gemm_fp8_edge_kernel, dim3(ceil_div(m, 16), ceil_div(n, 16)), dim3(256), 0,
                       s, m, n, k, Ap, ldak, Bp, ldbk, scales, C, ldc);
    SLA_LAUNCH_CHECK();
}

// SLA_SPLIT_LDS=0 in the environment keeps the split product on the
// LDS-free 128 x 128 kernel (for comparisons).
static bool split_lds_on()
{
    static const bool on = [] {
        const char *e = getenv("SLA_SPLIT_LDS");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <typename T16, int S>
static void gemm_split_launch(int m, int n, int k, const T16 *Ap, int ldak, size_t aplane,
                              const T16 *Bp, int ldbk, size_t bplane, const float *absmax,
                              float *C, int ldc, hipStream_t s,
                              unsigned long long *stamp = nullptr)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    const bool vec = k % 32 == 0 && ldak % 8 == 0 && ldbk % 8 == 0 && ldc % 4 == 0 &&
                     aplane % 8 == 0 && bplane % 8 == 0 && aligned16(Ap) && aligned16(Bp) &&
                     aligned16(C);
    constexpr int TM = split_tile_m(S);
    const int ml = (m / TM) * TM, nl = (n / SL_N) * SL_N;
    if (split_lds_on() && vec && ml > 0 && nl > 0) {
        // TM x 128 tiles through LDS; the borders by the kernels below
        static bool attr = false;
        if (!attr) {
            SLA_HIP(hipFuncSetAttribute(
                reinterpret_cast<const void *>(&gemm_split_lds_kernel<T16, S>),
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)split_lds_bytes(S)));
            attr = true;
        }
        hipLaunchKernelGGL((gemm_split_lds_kernel<T16, S>), dim3((ml / TM) * (nl / SL_N)),
                           dim3(256), split_lds_bytes(S), s, ml, nl, k, Ap, ldak, aplane, Bp, ldbk,
                           bplane, absmax, C, ldc, stamp);
        SLA_LAUNCH_CHECK();
        if (m > ml)
            gemm_split_launch<T16, S>(m - ml, n, k, Ap + (size_t)ml * ldak, ldak, aplane, Bp, ldbk,
                                      bplane, absmax, C + ml, ldc, s);
        if (n > nl)
            gemm_split_launch<T16, S>(ml, n - nl, k, Ap, ldak, aplane, Bp + (size_t)nl * ldbk, ldbk,
                                      bplane, absmax, C + (size_t)nl * ldc, ldc, s);
        return;
    }
    const int mf = (m / 128) * 128, nf = (n / 128) * 128;
    if (mf > 0 && nf > 0 && vec) {
        hipLaunchKernelGGL((gemm_split_nn_sub_direct_kernel<T16, S>), dim3((mf / 128) * (nf / 128)),
                           dim3(256), 0, s, mf, nf, k, Ap, ldak, aplane, Bp, ldbk, bplane, absmax,
                           C, ldc, stamp);
        SLA_LAUNCH_CHECK();
        if (m > mf)
            gemm_split_launch<T16, S>(m - mf, n, k, Ap + (size_t)mf * ldak, ldak, aplane, Bp, ldbk,
                                      bplane, absmax, C + mf, ldc, s);
        if (n > nf)
            gemm_split_launch<T16, S>(mf, n - nf, k, Ap, ldak, aplane, Bp + (size_t)nf * ldbk, ldbk,
                                      bplane, absmax, C + (size_t)nf * ldc, ldc, s);
        return;
    }
    hipLaunchKernelGGL((gemm_split_edge_kernel<T16, S>), dim3(ceil_div(m, 16), ceil_div(n, 16)),
                       dim3(256), 0, s, m, n, k, Ap, ldak, aplane, Bp, ldbk, bplane, absmax, C,
                       ldc);
    SLA_LAUNCH_CHECK();
}

void gemm_split_nn_sub(int split, int m, int n, int k, const void *Ap, int ldak, size_t aplane,
                       const void *Bp, int ldbk, size_t bplane, const float *absmax, float *C,
                       int ldc, hipStream_t s, unsigned long long *stamp)
{
    switch (split) {
    case SPLIT_BF16X3:
        gemm_split_launch<__bf16, 2>(m, n, k, static_cast<const __bf16 *>(Ap), ldak, aplane,
                                     static_cast<const __bf16 *>(Bp), ldbk, bplane, absmax, C,
                                     ldc, s, stamp);
        break;
    case SPLIT_BF16X6:
        gemm_split_launch<__bf16, 3>(m, n, k, static_cast<const __bf16 *>(Ap), ldak, aplane,
                                     static_cast<const __bf16 *>(Bp), ldbk, bplane, absmax, C,
                                     ldc, s, stamp);
        break;
    case SPLIT_FP16X3:
        gemm_split_launch<_Float16, 2>(m, n, k, static_cast<const _Float16 *>(Ap), ldak, aplane,
                                       static_cast<const _Float16 *>(Bp), ldbk, bplane, absmax, C,
                                       ldc, s, stamp);
        break;
    default: throw std::invalid_argument("split GEMM: unknown split");
    }
}

template <typename T16, int S>
static void pack_split_launch(bool rows, int d0, int k, int kpad, const float *A, int lda,
                              const float *absmax, void *P, size_t plane, hipStream_t s)
{
    T16 *p = static_cast<T16 *>(P);
    if (rows) {  // d0 = m
        if (d0 <= 0 || kpad <= 0) return;
        hipLaunchKernelGGL((pack_rows_split_kernel<T16, S>), dim3(ceil_div(d0, 32), ceil_div(kpad, 32)),
                           dim3(256), 0, s, d0, k, kpad, A, lda, absmax, p, plane);
    } else {     // d0 = n
        if (d0 <= 0 || kpad <= 0) return;
        const size_t total = (size_t)kpad * d0;
        hipLaunchKernelGGL((pack_cols_split_kernel<T16, S>), dim3((unsigned)((total + 255) / 256)),
                           dim3(256), 0, s, k, kpad, d0, A, lda, absmax, p, plane);
    }
    SLA_LAUNCH_CHECK();
}

void pack_split(int split, bool rows, int d0, int k, int kpad, const float *A, int lda,
                const float *absmax, void *P, size_t plane, hipStream_t s)
{
    switch (split) {
    case SPLIT_BF16X3: pack_split_launch<__bf16, 2>(rows, d0, k, kpad, A, lda, absmax, P, plane, s); break;
    case SPLIT_BF16X6: pack_split_launch<__bf16, 3>(rows, d0, k, kpad, A, lda, absmax, P, plane, s); break;
    case SPLIT_FP16X3: pack_split_launch<_Float16, 2>(rows, d0, k, kpad, A, lda, absmax, P, plane, s); break;
    default: throw std::invalid_argument("split packing: unknown split");
    }
}

void convert_f64_f32(int m, int n, const double *A, int lda, float *B, int ldb, int *overflow,
                     hipStream_t s)
{
    if (m <= 0 || n <= 0) return;
    const size_t total = (size_t)m * n;
    hipLaunchKernelGGL(convert_f64_f32_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       s, m, n, A, lda, B, ldb, overflow);
    SLA_LAUNCH_CHECK();
}

}  // namespace sla
