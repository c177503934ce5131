"""scalapack_amd -- ScaLAPACK's dense LU solve path (PDGESV = PDGETRF +
PDGETRS on a 1 x 1 BLACS grid) rebuilt for one AMD Instinct MI355X (gfx950).

The product is the native library libscalapack_amd.so (HIP kernels + a C-ABI
that exports the reference's Fortran entry points, include/scalapack_amd.h).
This package loads it and mirrors the reference interface for Python callers:

    scalapack_amd.api     PDGESV / PDGETRF / PDGETRS / PDGETF2 / PDLASWP on
                          host arrays, BLACS 1 x 1 context and DESCINIT
    scalapack_amd.device  the same solve on matrices already in HBM
    scalapack_amd.build   the in-tree build (hipcc --offload-arch=gfx950)
"""
from ._lib import LibraryMissing, load  # noqa: F401

__all__ = ["LibraryMissing", "load"]
