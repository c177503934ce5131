"""ASCII matrix I/O of the TOOLS shim (PDLAWRITE / PDLAREAD, host code, no
GPU) against the reference's own files: the reference built from source
wrote tests/golden/pdlawrite_9x6.txt (TOOLS/pdlawrite.f: list-directed
header, values in D30.18 down the columns, three-digit exponents without
the D) and read EXAMPLE/DSCAEXMAT.dat into tests/golden/pdlaread_dscaexmat.npy
(TOOLS/pdlaread.f).  Both are compared bit for bit."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from scalapack_amd import api
from scalapack_amd._lib import shim

GOLD = Path(__file__).resolve().parent / "golden"
_DAT = np.load(GOLD / "example_6x6.npz")


def _fortran_float(tok: str) -> float:
    """A D30.18 field: D exponent, or a three-digit exponent with no letter."""
    t = tok.strip().replace("D", "E")
    if "E" not in t:
        k = max(t.rfind("+"), t.rfind("-"))
        t = t[:k] + "E" + t[k:]
    return float(t)


def _w(path, m, n, a, ia, ja, desc):
    lib = shim()
    d = np.ascontiguousarray(desc, np.int32)
    b = str(path).encode()
    lib.pdlawrite_(ctypes.c_char_p(b), ctypes.byref(ctypes.c_int(m)), ctypes.byref(ctypes.c_int(n)),
                   ctypes.c_void_p(a.ctypes.data), ctypes.byref(ctypes.c_int(ia)),
                   ctypes.byref(ctypes.c_int(ja)), ctypes.c_void_p(d.ctypes.data),
                   ctypes.byref(ctypes.c_int(0)), ctypes.byref(ctypes.c_int(0)), None,
                   ctypes.c_size_t(len(b)))


def _r(path, a, desc):
    lib = shim()
    d = np.ascontiguousarray(desc, np.int32)
    b = str(path).encode()
    lib.pdlaread_(ctypes.c_char_p(b), ctypes.c_void_p(a.ctypes.data), ctypes.c_void_p(d.ctypes.data),
                  ctypes.byref(ctypes.c_int(0)), ctypes.byref(ctypes.c_int(0)), None,
                  ctypes.c_size_t(len(b)))


def test_pdlawrite_matches_reference_bytes(oracle, tmp_path):
    M, N, MB = 9, 6, 4
    g = oracle.matgen(M, N, 100)
    flat = g.ravel(order="F")
    for k, v in enumerate([0.0, 1.0, -0.5, 1e-300, -3e300, 123456789.0]):
        flat[(k * 7) % (M * N)] = v
    g = flat.reshape((M, N), order="F")
    ctx = api.blacs_gridinit(1, 1)
    desc, _ = api.descinit(M + 2, N + 1, MB, MB, 0, 0, ctx, M + 2)
    a = np.zeros((M + 2, N + 1), order="F")
    a[:M, :N] = g
    out = tmp_path / "w.txt"
    _w(out, M, N, a, 1, 1, desc)
    assert out.read_bytes() == (GOLD / "pdlawrite_9x6.txt").read_bytes()
    # a submatrix: IA = 2, JA = 2 writes A(2:.., 2:..)
    _w(out, 3, 2, a, 2, 2, desc)
    lines = out.read_text().splitlines()
    assert lines[0] == " 3 2" and len(lines) == 7
    assert [_fortran_float(x) for x in lines[1:]] == [a[1, 1], a[2, 1], a[3, 1], a[1, 2], a[2, 2], a[3, 2]]


def test_pdlaread_matches_reference(tmp_path):
    """The 6 x 6 example file (EXAMPLE/DSCAEXMAT.dat's format: "6 6", then one
    value per record with a D exponent) read into a 12 x 8 array: the same
    array as the reference's PDLAREAD produced."""
    ref = np.load(GOLD / "pdlaread_dscaexmat.npy", allow_pickle=False)
    a6 = _DAT["a"]
    src = tmp_path / "m.dat"
    with open(src, "w") as f:
        f.write("6 6\n")
        for v in a6.ravel(order="F"):
            f.write(f" {v:.4f}D+0\n")
    ctx = api.blacs_gridinit(1, 1)
    desc, _ = api.descinit(12, 8, 4, 4, 0, 0, ctx, 12)
    a = np.zeros((12, 8), order="F")
    _r(src, a, desc)
    assert np.array_equal(a, ref)
    assert np.array_equal(a[:6, :6], a6)


def test_roundtrip_is_exact(tmp_path):
    """D30.18 holds 18 significant digits, so write -> read returns every
    double bit for bit (17 digits suffice)."""
    rng = np.random.default_rng(4)
    g = np.asfortranarray(rng.standard_normal((13, 7)) * 10.0 ** rng.integers(-200, 200, (13, 7)))
    g[0, 0], g[1, 0] = 5e-324, -1.7976931348623157e308
    ctx = api.blacs_gridinit(1, 1)
    desc, _ = api.descinit(13, 7, 4, 4, 0, 0, ctx, 13)
    f = tmp_path / "rt.txt"
    _w(f, 13, 7, g, 1, 1, desc)
    b = np.zeros((13, 7), order="F")
    _r(f, b, desc)
    assert np.array_equal(b, g)
