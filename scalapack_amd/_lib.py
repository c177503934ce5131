"""Loader of the native libraries (ctypes).

The product path is the HIP library; there is no Python or CPU fallback.  If
the library is missing, every entry point raises LibraryMissing.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HEADERS = [ROOT / "include" / "scalapack_amd.h", ROOT / "include" / "sla_blacs1x1.h"]

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)


class LibraryMissing(RuntimeError):
    pass


class SolveInfo(ctypes.Structure):
    """sla_solve_info_t of include/scalapack_amd.h."""
    _fields_ = [("precision", ctypes.c_int), ("refinement", ctypes.c_int),
                ("iter", ctypes.c_int), ("outer", ctypes.c_int), ("inner", ctypes.c_int),
                ("fell_back", ctypes.c_int), ("anorm", ctypes.c_double),
                ("bwd_err", ctypes.c_double), ("sresid", ctypes.c_double),
                ("history", ctypes.c_double * 16), ("stop_rule", ctypes.c_int)]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["history"] = [v for v in self.history[: min(16, self.outer + 1)]]
        return d


_lib = None
_shim = None


def _declare(lib: ctypes.CDLL) -> None:
    vp, ip, i, d = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sz = ctypes.c_size_t
    sigs = {
        "pdgesv_": [vp] * 12,
        "pdgetrf_": [vp] * 8,
        "pdgetf2_": [vp] * 8,
        "pdgetrs_": [vp] * 13 + [sz],
        "pdlaswp_": [vp] * 10 + [sz, sz],
        "pdgecon_": [vp] * 13 + [sz],
        "pdgerfs_": [vp] * 27 + [sz],
        "pdgeequ_": [vp] * 12,
        "pdlaqge_": [vp] * 12 + [sz],
        "pdgesvx_": [vp] * 32 + [sz, sz, sz],
        "sla_create": [vp],
        "sla_destroy": [vp],
        "sla_set_stream": [vp, vp],
        "sla_set_param": [vp, ctypes.c_char_p, i],
        "sla_get_param": [vp, ctypes.c_char_p, ip],
        "sla_get_stats": [vp, i, vp, vp, vp],
        "sla_get_update_exec": [vp, vp, vp, vp],
        "sla_reset_stats": [vp],
        "sla_dgetrf": [vp, i, i, vp, i, vp, vp],
        "sla_dgetrs": [vp, i, i, vp, i, vp, vp, i],
        "sla_dgesv": [vp, i, i, vp, i, vp, vp, i, vp],
        "sla_dgemm_sub": [vp, i, i, i, vp, i, vp, i, vp, i],
        "sla_dxgesv": [vp, i, i, i, i, vp, i, vp, vp, i, vp, i, vp, vp],
        "sla_last_solve_info": [vp, vp],
        "sla_sgetrf": [vp, i, i, i, vp, i, vp, vp],
        "sla_gemm_lowp_sub": [vp, i, i, i, i, vp, i, vp, i, vp, i],
        "sla_sgetrs": [vp, i, i, vp, i, vp, vp, i],
        "sla_dresidual": [vp, i, i, vp, i, vp, i, vp, i, vp, i, vp],
        "sla_dgetrs_trans": [vp, ctypes.c_char, i, i, vp, i, vp, vp, i],
        "sla_dlange": [vp, ctypes.c_char, i, i, vp, i, vp],
        "sla_dgecon": [vp, ctypes.c_char, i, vp, i, d, vp],
        "sla_dgerfs": [vp, ctypes.c_char, i, i, vp, i, vp, i, vp, vp, i, vp, i, vp, vp],
        "sla_gesv_batched": [vp, i, i, vp, vp, vp, vp, vp, vp, vp, vp],
        "sla_gesv_batched_tune": [vp, i, i, vp, vp, vp, vp, vp],
        "sla_selector_features": [vp, i, vp, i, vp],
        "sla_dgeequ": [vp, i, i, vp, i, vp, vp, vp, vp, vp, vp],
        "sla_dlaqge": [vp, i, i, vp, i, vp, vp, d, d, d, vp],
        "sla_dgesvx": [vp, ctypes.c_char, ctypes.c_char, i, i, vp, i, vp, i, vp, vp, vp, vp, vp,
                       i, vp, i, vp, vp, vp, vp],
        "sla_mlp_grad": [vp, i, vp, vp, vp, vp, vp, vp],
        "sla_mlp_train": [vp, i, vp, vp, vp, vp, i, ctypes.c_float, vp],
        "sla_mlp_predict": [vp, i, vp, vp, vp, vp],
        "sla_selector_set_model": [vp, vp],
        "sla_dxgesv_auto": [vp, i, i, vp, i, vp, vp, i, vp, i, vp, vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = None if name.endswith("_") else ctypes.c_int
    lib.pdlange_.argtypes = [vp] * 8 + [sz]
    lib.pdlange_.restype = ctypes.c_double
    lib.sla_last_error.argtypes = []
    lib.sla_last_error.restype = ctypes.c_char_p


def load() -> ctypes.CDLL:
    """Load libscalapack_amd.so (and its 1 x 1 BLACS shim) once."""
    global _lib, _shim
    if _lib is not None:
        return _lib
    so = PKG / "libscalapack_amd.so"
    shim = PKG / "libsla_blacs1x1.so"
    if not so.exists() or not shim.exists():
        raise LibraryMissing(
            f"{so} is not built; run `python -m scalapack_amd.build` (hipcc, gfx950)")
    _shim = ctypes.CDLL(str(shim), mode=ctypes.RTLD_GLOBAL)
    _lib = ctypes.CDLL(str(so), mode=ctypes.RTLD_GLOBAL)
    _declare(_lib)
    return _lib


def shim() -> ctypes.CDLL:
    load()
    return _shim


def declared_symbols() -> dict[str, list[str]]:
    """Function names declared by each public header (for the export check)."""
    out = {}
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
        out[h.name] = sorted(set(n for n in names if n not in ("if", "while", "return")))
    return out


def last_error() -> str:
    return load().sla_last_error().decode()
