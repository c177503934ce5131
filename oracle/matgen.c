/*
 * matgen.c -- the reference tester's matrix generator, restated
 * (TEST INFRASTRUCTURE ONLY, see oracle.h).
 *
 * PDMATGEN (TESTING/traditional/LIN/pdmatgen.f:448-510) draws every entry of a
 * general matrix from one 31-bit linear congruential stream
 *     X(k+1) = (1103515245 * X(k) + 12345) mod 2^31,   X(0) = ISEED
 * (multiplier 16838 * 2^16 + 20077, increment 12345: pdmatgen.f:116-118;
 *  LMUL/LADD keep 16 + 15 bits: pmatgeninc.f:28-29,67-74) and sets
 *     A(i, j) = 1 - 2 * X((j-1)*M + i) / 2^31      (1-based i, j)
 * (PDRAND, pmatgeninc.f:311-314, first value X(1) after SETRAN).  The row and
 * column jumps of pdmatgen.f only reposition the same global stream, so the
 * value depends on the global index alone, whatever the process grid.
 * Positions are reached here by squaring the affine map, not by stepping.
 */
#include "oracle.h"

#include <stddef.h>
#include <stdint.h>

#define LCG_A 1103515245ull
#define LCG_C 12345ull
#define LCG_MASK 0x7fffffffull

/* f^k(x) = A x + C (mod 2^31) for f(x) = LCG_A x + LCG_C. */
static void lcg_jump(uint64_t k, uint64_t *A, uint64_t *C)
{
    uint64_t ra = 1, rc = 0;               /* identity */
    uint64_t pa = LCG_A, pc = LCG_C;       /* f^(2^bit) */
    while (k) {
        if (k & 1) {                       /* r = p o r */
            rc = (pa * rc + pc) & LCG_MASK;
            ra = (pa * ra) & LCG_MASK;
        }
        pc = (pa * pc + pc) & LCG_MASK;    /* p = p o p */
        pa = (pa * pa) & LCG_MASK;
        k >>= 1;
    }
    *A = ra;
    *C = rc;
}

unsigned oracle_lcg_state(int seed, long long k)
{
    uint64_t A, C;
    lcg_jump((uint64_t)k, &A, &C);
    return (unsigned)((A * ((uint64_t)(uint32_t)seed & LCG_MASK) + C) & LCG_MASK);
}

void oracle_matgen(int M, int N, int seed, int i0, int j0, int mloc, int nloc,
                   double *out, int ldo)
{
    (void)N;
    uint64_t a1, c1;
    lcg_jump(1, &a1, &c1);
    for (int j = 0; j < nloc; j++) {
        long long k = (long long)(j0 + j) * M + i0 + 1;
        uint64_t x = oracle_lcg_state(seed, k);
        double *col = out + (size_t)j * ldo;
        for (int i = 0; i < mloc; i++) {
            col[i] = 1.0 - 2.0 * ((double)x / 2147483648.0);
            x = (a1 * x + c1) & LCG_MASK;
        }
    }
}
