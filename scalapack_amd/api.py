"""The reference's LU-path interface, callable from Python on numpy arrays.

Each function takes the same arguments, with the same meaning, as the
Fortran routine it mirrors and returns INFO, so a test reads like the
reference's own testers.  Arrays are the *local* arrays of a 1 x 1 BLACS
grid: column-major (Fortran-ordered) float64 with leading dimension
DESC(LLD_); IA/JA/IB/JB are 1-based; IPIV is int32 of length >= LOCr(M_A) +
MB_A and receives 1-based global row numbers.  All work runs in
libscalapack_amd.so through its C-ABI (the symbols a Fortran program links).

    ictxt = blacs_gridinit(1, 1)
    desca = descinit(n, n, nb, nb, 0, 0, ictxt, lld)
    info = pdgesv(n, nrhs, a, 1, 1, desca, ipiv, b, 1, 1, descb)

References: SRC/pdgesv.f:1-2, SRC/pdgetrf.f:1, SRC/pdgetrs.f:1-2,
SRC/pdgetf2.f:1, SRC/pdlaswp.f:1-2, SRC/pdlange.f:1-2, SRC/pdgecon.f:1-3,
SRC/pdgerfs.f:1-4, TOOLS/descinit.f, TOOLS/numroc.f.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import load, shim

DTYPE_, CTXT_, M_, N_, MB_, NB_, RSRC_, CSRC_, LLD_ = range(9)


def _i(v: int) -> ctypes.c_int:
    return ctypes.c_int(int(v))


def _ref(x):
    return ctypes.byref(x)


def _arr(a: np.ndarray, dtype) -> ctypes.c_void_p:
    if a.dtype != dtype:
        raise TypeError(f"expected {np.dtype(dtype).name} array, got {a.dtype}")
    if a.ndim == 2 and not a.flags.f_contiguous:
        raise TypeError("matrices must be Fortran-ordered (column-major)")
    if a.ndim == 1 and not a.flags.c_contiguous:
        raise TypeError("vectors must be contiguous")
    return ctypes.c_void_p(a.ctypes.data)


def _desc(d) -> np.ndarray:
    d = np.ascontiguousarray(d, dtype=np.int32)
    if d.shape != (9,):
        raise ValueError("a descriptor has 9 integer entries")
    return d


# ---------------------------------------------------------------- BLACS / TOOLS
def blacs_pinfo() -> tuple[int, int]:
    me, n = ctypes.c_int(), ctypes.c_int()
    shim().blacs_pinfo_(_ref(me), _ref(n))
    return me.value, n.value


def blacs_gridinit(nprow: int = 1, npcol: int = 1) -> int:
    s = shim()
    ctxt = ctypes.c_int()
    s.blacs_get_(_ref(_i(-1)), _ref(_i(0)), _ref(ctxt))
    s.blacs_gridinit_(_ref(ctxt), b"Row-major", _ref(_i(nprow)), _ref(_i(npcol)),
                      ctypes.c_size_t(9))
    return ctxt.value


def blacs_gridinfo(ictxt: int) -> tuple[int, int, int, int]:
    v = [ctypes.c_int() for _ in range(4)]
    shim().blacs_gridinfo_(_ref(_i(ictxt)), *map(_ref, v))
    return tuple(x.value for x in v)


def blacs_gridexit(ictxt: int) -> None:
    shim().blacs_gridexit_(_ref(_i(ictxt)))


def numroc(n: int, nb: int, iproc: int, isrcproc: int, nprocs: int) -> int:
    f = shim().numroc_
    f.restype = ctypes.c_int
    return f(_ref(_i(n)), _ref(_i(nb)), _ref(_i(iproc)), _ref(_i(isrcproc)), _ref(_i(nprocs)))


def descinit(m, n, mb, nb, irsrc, icsrc, ictxt, lld) -> tuple[np.ndarray, int]:
    desc = np.zeros(9, np.int32)
    info = ctypes.c_int()
    shim().descinit_(ctypes.c_void_p(desc.ctypes.data), *(_ref(_i(v)) for v in
                     (m, n, mb, nb, irsrc, icsrc, ictxt, lld)), _ref(info))
    return desc, info.value


# ------------------------------------------------------------------ LU path
def pdgetrf(m, n, a, ia, ja, desca, ipiv) -> int:
    """PDGETRF (SRC/pdgetrf.f): A(IA:IA+M-1, JA:JA+N-1) = P L U in place."""
    d = _desc(desca)
    info = ctypes.c_int()
    load().pdgetrf_(_ref(_i(m)), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                    ctypes.c_void_p(d.ctypes.data), _arr(ipiv, np.int32), _ref(info))
    return info.value


def pdgetf2(m, n, a, ia, ja, desca, ipiv) -> int:
    """PDGETF2 (SRC/pdgetf2.f): unblocked panel factorization, N <= NB."""
    d = _desc(desca)
    info = ctypes.c_int()
    load().pdgetf2_(_ref(_i(m)), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                    ctypes.c_void_p(d.ctypes.data), _arr(ipiv, np.int32), _ref(info))
    return info.value


def pdgetrs(trans, n, nrhs, a, ia, ja, desca, ipiv, b, ib, jb, descb) -> int:
    """PDGETRS (SRC/pdgetrs.f): solve with the factors of PDGETRF."""
    da, db = _desc(desca), _desc(descb)
    info = ctypes.c_int()
    t = ctypes.c_char(trans.encode()[:1] if isinstance(trans, str) else trans)
    load().pdgetrs_(_ref(t), _ref(_i(n)), _ref(_i(nrhs)), _arr(a, np.float64), _ref(_i(ia)),
                    _ref(_i(ja)), ctypes.c_void_p(da.ctypes.data), _arr(ipiv, np.int32),
                    _arr(b, np.float64), _ref(_i(ib)), _ref(_i(jb)),
                    ctypes.c_void_p(db.ctypes.data), _ref(info), ctypes.c_size_t(1))
    return info.value


def pdgesv(n, nrhs, a, ia, ja, desca, ipiv, b, ib, jb, descb) -> int:
    """PDGESV (SRC/pdgesv.f): A X = B; A <- L\\U, B <- X."""
    da, db = _desc(desca), _desc(descb)
    info = ctypes.c_int()
    load().pdgesv_(_ref(_i(n)), _ref(_i(nrhs)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                   ctypes.c_void_p(da.ctypes.data), _arr(ipiv, np.int32), _arr(b, np.float64),
                   _ref(_i(ib)), _ref(_i(jb)), ctypes.c_void_p(db.ctypes.data), _ref(info))
    return info.value


def pdlange(norm, m, n, a, ia, ja, desca, work=None) -> float:
    """PDLANGE (SRC/pdlange.f): 'M', '1'/'O', 'I' or 'F'/'E' norm of sub(A)."""
    d = _desc(desca)
    w = np.zeros(1) if work is None else work
    c = ctypes.c_char(norm.encode()[:1])
    return load().pdlange_(_ref(c), _ref(_i(m)), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)),
                           _ref(_i(ja)), ctypes.c_void_p(d.ctypes.data), _arr(w, np.float64),
                           ctypes.c_size_t(1))


def pdgecon(norm, n, a, ia, ja, desca, anorm, work, lwork, iwork, liwork) -> tuple[float, int]:
    """PDGECON (SRC/pdgecon.f): returns (RCOND, INFO).  WORK(1) / IWORK(1)
    receive the minimal LWORK / LIWORK (LWORK = -1 or LIWORK = -1 queries)."""
    d = _desc(desca)
    rcond, info = ctypes.c_double(), ctypes.c_int()
    c = ctypes.c_char(norm.encode()[:1])
    load().pdgecon_(_ref(c), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                    ctypes.c_void_p(d.ctypes.data), _ref(ctypes.c_double(anorm)), _ref(rcond),
                    _arr(work, np.float64), _ref(_i(lwork)), _arr(iwork, np.int32),
                    _ref(_i(liwork)), _ref(info), ctypes.c_size_t(1))
    return rcond.value, info.value


def pdgerfs(trans, n, nrhs, a, ia, ja, desca, af, iaf, jaf, descaf, ipiv, b, ib, jb, descb, x,
            ix, jx, descx, ferr, berr, work, lwork, iwork, liwork) -> int:
    """PDGERFS (SRC/pdgerfs.f): refines X in place, FERR / BERR per column."""
    da, daf, db, dx = _desc(desca), _desc(descaf), _desc(descb), _desc(descx)
    info = ctypes.c_int()
    t = ctypes.c_char(trans.encode()[:1])
    vp = ctypes.c_void_p
    load().pdgerfs_(_ref(t), _ref(_i(n)), _ref(_i(nrhs)), _arr(a, np.float64), _ref(_i(ia)),
                    _ref(_i(ja)), vp(da.ctypes.data), _arr(af, np.float64), _ref(_i(iaf)),
                    _ref(_i(jaf)), vp(daf.ctypes.data), _arr(ipiv, np.int32),
                    _arr(b, np.float64), _ref(_i(ib)), _ref(_i(jb)), vp(db.ctypes.data),
                    _arr(x, np.float64), _ref(_i(ix)), _ref(_i(jx)), vp(dx.ctypes.data),
                    _arr(ferr, np.float64), _arr(berr, np.float64), _arr(work, np.float64),
                    _ref(_i(lwork)), _arr(iwork, np.int32), _ref(_i(liwork)), _ref(info),
                    ctypes.c_size_t(1))
    return info.value


def pdgeequ(m, n, a, ia, ja, desca, r, c) -> tuple[float, float, float, int]:
    """PDGEEQU (SRC/pdgeequ.f): returns (ROWCND, COLCND, AMAX, INFO)."""
    d = _desc(desca)
    rc, cc, am, info = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    load().pdgeequ_(_ref(_i(m)), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                    ctypes.c_void_p(d.ctypes.data), _arr(r, np.float64), _arr(c, np.float64),
                    _ref(rc), _ref(cc), _ref(am), _ref(info))
    return rc.value, cc.value, am.value, info.value


def pdlaqge(m, n, a, ia, ja, desca, r, c, rowcnd, colcnd, amax) -> str:
    """PDLAQGE (SRC/pdlaqge.f): returns EQUED."""
    d = _desc(desca)
    e = ctypes.create_string_buffer(1)
    load().pdlaqge_(_ref(_i(m)), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)), _ref(_i(ja)),
                    ctypes.c_void_p(d.ctypes.data), _arr(r, np.float64), _arr(c, np.float64),
                    _ref(ctypes.c_double(rowcnd)), _ref(ctypes.c_double(colcnd)),
                    _ref(ctypes.c_double(amax)), e, ctypes.c_size_t(1))
    return e.raw.decode()


def pdgesvx(fact, trans, n, nrhs, a, ia, ja, desca, af, iaf, jaf, descaf, ipiv, equed, r, c, b,
            ib, jb, descb, x, ix, jx, descx, ferr, berr, work, lwork, iwork,
            liwork) -> tuple[int, str, float]:
    """PDGESVX (SRC/pdgesvx.f): returns (INFO, EQUED, RCOND)."""
    da, daf, db, dx = _desc(desca), _desc(descaf), _desc(descb), _desc(descx)
    e = ctypes.create_string_buffer(equed.encode()[:1], 1)
    rc, info = ctypes.c_double(), ctypes.c_int()
    vp = ctypes.c_void_p
    f = ctypes.c_char(fact.encode()[:1])
    t = ctypes.c_char(trans.encode()[:1])
    load().pdgesvx_(_ref(f), _ref(t), _ref(_i(n)), _ref(_i(nrhs)), _arr(a, np.float64),
                    _ref(_i(ia)), _ref(_i(ja)), vp(da.ctypes.data), _arr(af, np.float64),
                    _ref(_i(iaf)), _ref(_i(jaf)), vp(daf.ctypes.data), _arr(ipiv, np.int32), e,
                    _arr(r, np.float64), _arr(c, np.float64), _arr(b, np.float64), _ref(_i(ib)),
                    _ref(_i(jb)), vp(db.ctypes.data), _arr(x, np.float64), _ref(_i(ix)),
                    _ref(_i(jx)), vp(dx.ctypes.data), _ref(rc), _arr(ferr, np.float64),
                    _arr(berr, np.float64), _arr(work, np.float64), _ref(_i(lwork)),
                    _arr(iwork, np.int32), _ref(_i(liwork)), _ref(info), ctypes.c_size_t(1),
                    ctypes.c_size_t(1), ctypes.c_size_t(1))
    return info.value, e.raw.decode(), rc.value


def pdlaswp(direc, rowcol, n, a, ia, ja, desca, k1, k2, ipiv) -> None:
    """PDLASWP (SRC/pdlaswp.f): row or column interchanges K1..K2."""
    d = _desc(desca)
    dc = ctypes.c_char(direc.encode()[:1])
    rc = ctypes.c_char(rowcol.encode()[:1])
    load().pdlaswp_(_ref(dc), _ref(rc), _ref(_i(n)), _arr(a, np.float64), _ref(_i(ia)),
                    _ref(_i(ja)), ctypes.c_void_p(d.ctypes.data), _ref(_i(k1)), _ref(_i(k2)),
                    _arr(ipiv, np.int32), ctypes.c_size_t(1), ctypes.c_size_t(1))
