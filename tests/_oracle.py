"""ctypes access to oracle/liboracle.so (the CPU restatement of the
reference's 1 x 1 LU path, see oracle/oracle.h).  Test infrastructure: only
tests, the smoke check and the benchmark's CPU-baseline leg use it, and only
as a checker or a timed baseline."""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
SO = ROOT / "oracle" / "liboracle.so"

# The oracle calls MKL; fix its code path so results do not depend on the
# host CPU's instruction set more than necessary.
os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True,
                           capture_output=True)
        _lib = ctypes.CDLL(str(SO))
        _lib.oracle_sresid.restype = ctypes.c_double
        _lib.oracle_fresid.restype = ctypes.c_double
        _lib.oracle_norm_inf.restype = ctypes.c_double
        _lib.oracle_lcg_state.restype = ctypes.c_uint
        _lib.oracle_lcg_state.argtypes = [ctypes.c_int, ctypes.c_longlong]
        _lib.oracle_pdlange.restype = ctypes.c_double
        _lib.oracle_pdlange.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p]
        _lib.oracle_pdgecon.restype = ctypes.c_double
        _lib.oracle_pdgecon.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_double]
    return _lib


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def desc(m, n, nb, lld=None) -> np.ndarray:
    return np.array([1, 0, m, n, nb, nb, 0, 0, lld if lld is not None else max(1, m)], np.int32)


def matgen(M: int, N: int, seed: int = 100, i0: int = 0, j0: int = 0, mloc=None, nloc=None):
    """PDMATGEN's global matrix (or the block starting at (i0, j0))."""
    mloc = M if mloc is None else mloc
    nloc = N if nloc is None else nloc
    a = np.zeros((mloc, nloc), order="F")
    lib().oracle_matgen(M, N, seed, i0, j0, mloc, nloc, _p(a), max(1, mloc))
    return a


def lcg_state(seed: int, k: int) -> int:
    return lib().oracle_lcg_state(seed, k)


def pdgetf2(a: np.ndarray, nb: int, m: int, n: int, ia=1, ja=1):
    """PDGETF2 on a(ia:ia+m-1, ja:ja+n-1) (Fortran-ordered, in place), column
    by column as SRC/pdgetf2.f:206-237; returns (ipiv of those rows, info)."""
    M, N = a.shape
    d = desc(M, N, nb, M)
    ipiv = np.zeros(M + nb, np.int32)
    info = lib().oracle_pdgetf2(m, n, _p(a), ia, ja, _p(d), _p(ipiv), None)
    mn = min(m, n)
    return ipiv[ia - 1: ia - 1 + mn].copy(), info


def pdgetrf(a: np.ndarray, nb: int, m=None, n=None, ia=1, ja=1, with_gap=False):
    """Factor a (Fortran-ordered, modified in place); returns (ipiv, info[, gap])."""
    M, N = a.shape
    m = M - ia + 1 if m is None else m
    n = N - ja + 1 if n is None else n
    d = desc(M, N, nb, M)
    ipiv = np.zeros(M + nb, np.int32)
    gap = np.zeros(max(1, min(m, n)), np.float64)
    info = lib().oracle_pdgetrf(m, n, _p(a), ia, ja, _p(d), _p(ipiv),
                                _p(gap) if with_gap else None)
    mn = min(m, n)
    piv = ipiv[ia - 1: ia - 1 + mn].copy()
    return (piv, info, gap[:mn]) if with_gap else (piv, info)


def pdgetrs(lu: np.ndarray, ipiv_full: np.ndarray, b: np.ndarray, nb: int) -> np.ndarray:
    """Solve with the oracle's factors; ipiv_full indexed by global row."""
    N = lu.shape[0]
    nrhs = b.shape[1]
    x = np.asfortranarray(b.copy())
    da, db = desc(N, N, nb, N), desc(N, nrhs, nb, N)
    piv = np.zeros(N + nb, np.int32)
    piv[: len(ipiv_full)] = ipiv_full
    lib().oracle_pdgetrs(ctypes.c_char(b"N"), N, nrhs, _p(lu), 1, 1, _p(da), _p(piv), _p(x), 1,
                         1, _p(db))
    return x


def pdlaswp(direc: str, a: np.ndarray, ja: int, n: int, k1: int, k2: int, ipiv: np.ndarray):
    M = a.shape[0]
    d = desc(M, a.shape[1], 1, M)
    lib().oracle_pdlaswp(ctypes.c_char(direc.encode()), n, _p(a), 1, ja, _p(d), k1, k2,
                         _p(np.ascontiguousarray(ipiv, np.int32)))


def sresid(a0: np.ndarray, x: np.ndarray, b: np.ndarray) -> float:
    n, nrhs = x.shape
    a0 = np.asfortranarray(a0)
    x = np.asfortranarray(x)
    b = np.asfortranarray(b)
    return lib().oracle_sresid(n, nrhs, _p(a0), a0.shape[0], _p(x), x.shape[0], _p(b),
                               b.shape[0])


def fresid(a0: np.ndarray, lu: np.ndarray, ipiv: np.ndarray) -> float:
    m, n = a0.shape
    a0 = np.asfortranarray(a0)
    lu = np.asfortranarray(lu)
    return lib().oracle_fresid(m, n, _p(a0), m, _p(lu), lu.shape[0],
                               _p(np.ascontiguousarray(ipiv, np.int32)))


def pdlange(norm: str, a: np.ndarray, m=None, n=None, ia=1, ja=1) -> float:
    """PDLANGE of a(ia:ia+m-1, ja:ja+n-1)."""
    a = np.asfortranarray(a)
    M, N = a.shape
    m = M - ia + 1 if m is None else m
    n = N - ja + 1 if n is None else n
    return lib().oracle_pdlange(norm.encode()[:1], m, n, _p(a), ia, ja, _p(desc(M, N, 1, M)))


def pdgecon(norm: str, lu: np.ndarray, anorm: float) -> float:
    """PDGECON's RCOND from L\\U factors (IA = JA = 1)."""
    lu = np.asfortranarray(lu)
    n = lu.shape[0]
    return lib().oracle_pdgecon(norm.encode()[:1], n, _p(lu), 1, 1, _p(desc(n, n, 1, n)), anorm)


def pdgetrs_t(lu: np.ndarray, ipiv: np.ndarray, b: np.ndarray) -> np.ndarray:
    """A^T X = B with the oracle's factors (PDGETRS TRANS = 'T')."""
    n = lu.shape[0]
    x = np.asfortranarray(b.copy())
    nrhs = x.shape[1]
    lib().oracle_pdgetrs_t(n, nrhs, _p(np.asfortranarray(lu)), 1, 1, _p(desc(n, n, 1, n)),
                           _p(np.ascontiguousarray(ipiv, np.int32)), _p(x), 1, 1,
                           _p(desc(n, nrhs, 1, n)))
    return x


def pdgerfs(trans: str, a: np.ndarray, lu: np.ndarray, ipiv: np.ndarray, b: np.ndarray,
            x0: np.ndarray):
    """PDGERFS from the start x0; returns (x, ferr, berr)."""
    a, lu, b = (np.asfortranarray(v) for v in (a, lu, b))
    x = np.asfortranarray(x0.copy())
    n, nrhs = b.shape
    ferr, berr = np.zeros(nrhs), np.zeros(nrhs)
    lib().oracle_pdgerfs(ctypes.c_char(trans.encode()[:1]), n, nrhs, _p(a), n, _p(lu), n,
                         _p(np.ascontiguousarray(ipiv, np.int32)), _p(b), n, _p(x), n,
                         _p(ferr), _p(berr))
    return x, ferr, berr


def pdgeequ(a: np.ndarray):
    """(R, C, ROWCND, COLCND, AMAX, INFO) of PDGEEQU."""
    a = np.asfortranarray(a)
    m, n = a.shape
    r, c = np.zeros(max(1, m)), np.zeros(max(1, n))
    rc, cc, am = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    info = lib().oracle_pdgeequ(m, n, _p(a), max(1, m), _p(r), _p(c), ctypes.byref(rc),
                                ctypes.byref(cc), ctypes.byref(am))
    return r[:m], c[:n], rc.value, cc.value, am.value, info


def pdlaqge(a: np.ndarray, r, c, rowcnd, colcnd, amax) -> str:
    """Scales a in place; returns EQUED."""
    lib().oracle_pdlaqge.restype = ctypes.c_char
    m, n = a.shape
    e = lib().oracle_pdlaqge(m, n, _p(a), max(1, m), _p(np.ascontiguousarray(r, np.float64)),
                             _p(np.ascontiguousarray(c, np.float64)), ctypes.c_double(rowcnd),
                             ctypes.c_double(colcnd), ctypes.c_double(amax))
    return e.decode()


def pdgesvx(fact: str, trans: str, a: np.ndarray, b: np.ndarray, nb: int):
    """PDGESVX with FACT 'E' or 'N' on copies of a and b; returns a dict."""
    a = np.asfortranarray(a.copy())
    b = np.asfortranarray(b.copy())
    n, nrhs = b.shape
    af, x = np.zeros((n, n), order="F"), np.zeros((n, nrhs), order="F")
    ipiv = np.zeros(n + nb, np.int32)
    r, c = np.zeros(n), np.zeros(n)
    ferr, berr = np.zeros(nrhs), np.zeros(nrhs)
    eq = ctypes.create_string_buffer(1)
    rcond = ctypes.c_double()
    info = lib().oracle_pdgesvx(ctypes.c_char(fact.encode()), ctypes.c_char(trans.encode()), n,
                                nrhs, nb, _p(a), _p(af), _p(ipiv), eq, _p(r), _p(c), _p(b),
                                _p(x), ctypes.byref(rcond), _p(ferr), _p(berr))
    return dict(info=info, equed=eq.raw.decode(), R=r, C=c, X=x, rcond=rcond.value, ferr=ferr,
                berr=berr, B=b, A=a, AF=af, ipiv=ipiv[:n])
