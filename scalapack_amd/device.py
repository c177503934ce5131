"""Device-resident interface (matrices already in HBM).

Handle wraps the sla_* C-ABI of include/scalapack_amd.h.  Tensors are torch
CUDA (HIP) tensors used as plain device memory: a column-major m x n matrix
with leading dimension lda is a tensor whose storage holds lda * n values,
e.g. the transpose view of a contiguous (n, lda) tensor.
"""
from __future__ import annotations

import ctypes

from ._lib import SolveInfo, last_error, load

STAT_CLASSES = ("panel", "laswp", "trsm", "gemm", "trsv", "other", "gemm_panel")

# factorization precisions and refinements (SLA_PREC_*, SLA_REFINE_*)
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2, "fp8": 3, "bf16x3": 4, "bf16x6": 5, "fp16x3": 6}
REFINEMENTS = {"ir": 0, "gmres": 1}
# what ended a mixed-precision refinement (SLA_STOP_*)
STOP_RULES = ("converged", "stall_accepted", "fallback", "failed")


class SlaError(RuntimeError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        raise SlaError(last_error())


def _ptr(t) -> ctypes.c_void_p:
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    if not t.is_cuda:
        raise TypeError("device interface: tensor must live on the GPU")
    return ctypes.c_void_p(t.data_ptr())


class Handle:
    """One library handle (streams, workspace, tuning) on the current device.

    The handle's work is ordered on one HIP stream.  By default that is
    torch's current stream, so tensors produced by torch are ready when a call
    starts; pass stream=None to keep the handle's own stream (then the caller
    synchronises its producers first)."""

    def __init__(self, stream="torch", **params):
        self._lib = load()
        h = ctypes.c_void_p()
        _check(self._lib.sla_create(ctypes.byref(h)))
        self._h = h
        if stream == "torch":
            import torch
            stream = torch.cuda.current_stream()
        if stream is not None:
            self.set_stream(stream)
        for k, v in params.items():
            self.set(k, v)

    def close(self) -> None:
        if self._h:
            self._lib.sla_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, name: str, value: int) -> None:
        _check(self._lib.sla_set_param(self._h, name.encode(), int(value)))

    def get(self, name: str) -> int:
        v = ctypes.c_int()
        _check(self._lib.sla_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_stream(self, stream) -> None:
        """stream: a torch.cuda.Stream or a raw hipStream_t int (0: HIP's
        default stream)."""
        raw = stream if isinstance(stream, int) else stream.cuda_stream
        _check(self._lib.sla_set_stream(self._h, ctypes.c_void_p(raw)))

    def stats(self) -> dict[str, dict[str, float]]:
        out = {}
        for i, name in enumerate(STAT_CLASSES):
            ms, work, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
            _check(self._lib.sla_get_stats(self._h, i, ctypes.byref(ms), ctypes.byref(work),
                                           ctypes.byref(n)))
            out[name] = {"ms": ms.value, "work": work.value, "launches": n.value}
        return out

    def update_exec(self) -> dict[str, float]:
        """Trailing-update launches as the kernels measure their own
        execution (first workgroup start to last workgroup end)."""
        ms, work, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        _check(self._lib.sla_get_update_exec(self._h, ctypes.byref(ms), ctypes.byref(work),
                                             ctypes.byref(n)))
        return {"ms": ms.value, "work": work.value, "launches": n.value}

    def reset_stats(self) -> None:
        _check(self._lib.sla_reset_stats(self._h))

    # ------------------------------------------------------------------ FP64
    def dgetrf(self, m: int, n: int, a, lda: int, ipiv) -> int:
        info = ctypes.c_int()
        _check(self._lib.sla_dgetrf(self._h, m, n, _ptr(a), lda, _ptr(ipiv), ctypes.byref(info)))
        return info.value

    def dgetrs(self, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int) -> None:
        _check(self._lib.sla_dgetrs(self._h, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb))

    def dgemm_sub(self, m: int, n: int, k: int, a, lda: int, b, ldb: int, c, ldc: int) -> None:
        """C -= A B (the trailing-update kernel alone)."""
        _check(self._lib.sla_dgemm_sub(self._h, m, n, k, _ptr(a), lda, _ptr(b), ldb, _ptr(c), ldc))

    def dgesv(self, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int) -> int:
        info = ctypes.c_int()
        _check(self._lib.sla_dgesv(self._h, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb,
                                   ctypes.byref(info)))
        return info.value

    def dgetrs_trans(self, trans: str, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int) -> None:
        """op(A) X = B with the factors of dgetrf; trans 'N', 'T' or 'C'."""
        _check(self._lib.sla_dgetrs_trans(self._h, trans.encode()[:1], n, nrhs, _ptr(a), lda,
                                          _ptr(ipiv), _ptr(b), ldb))

    def dlange(self, norm: str, m: int, n: int, a, lda: int) -> float:
        """PDLANGE: 'M', '1'/'O', 'I' or 'F'/'E' norm of an m x n matrix."""
        out = ctypes.c_double()
        _check(self._lib.sla_dlange(self._h, norm.encode()[:1], m, n, _ptr(a), lda,
                                    ctypes.byref(out)))
        return out.value

    def dgeequ(self, m: int, n: int, a, lda: int, r, c) -> tuple[float, float, float, int]:
        """PDGEEQU: scales into r, c (device); returns (rowcnd, colcnd, amax, info)."""
        rc, cc, am, info = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(self._lib.sla_dgeequ(self._h, m, n, _ptr(a), lda, _ptr(r), _ptr(c), ctypes.byref(rc),
                                    ctypes.byref(cc), ctypes.byref(am), ctypes.byref(info)))
        return rc.value, cc.value, am.value, info.value

    def dlaqge(self, m: int, n: int, a, lda: int, r, c, rowcnd, colcnd, amax) -> str:
        e = ctypes.create_string_buffer(1)
        _check(self._lib.sla_dlaqge(self._h, m, n, _ptr(a), lda, _ptr(r), _ptr(c), float(rowcnd),
                                    float(colcnd), float(amax), e))
        return e.raw.decode()

    def dgesvx(self, fact: str, trans: str, n: int, nrhs: int, a, lda: int, af, ldaf: int, ipiv,
               r, c, b, ldb: int, x, ldx: int, equed: str = "N"):
        """PDGESVX on device data: (info, equed, rcond, ferr, berr)."""
        e = ctypes.create_string_buffer(equed.encode()[:1], 1)
        rc, info = ctypes.c_double(), ctypes.c_int()
        fe = (ctypes.c_double * max(1, nrhs))()
        be = (ctypes.c_double * max(1, nrhs))()
        _check(self._lib.sla_dgesvx(self._h, fact.encode()[:1], trans.encode()[:1], n, nrhs, _ptr(a),
                                    lda, _ptr(af), ldaf, _ptr(ipiv), e, _ptr(r), _ptr(c), _ptr(b),
                                    ldb, _ptr(x), ldx, ctypes.byref(rc), fe, be, ctypes.byref(info)))
        return info.value, e.raw.decode(), rc.value, list(fe[:nrhs]), list(be[:nrhs])

    def dgecon(self, norm: str, n: int, lu, lda: int, anorm: float) -> float:
        """PDGECON: reciprocal condition estimate from the factors of dgetrf."""
        out = ctypes.c_double()
        _check(self._lib.sla_dgecon(self._h, norm.encode()[:1], n, _ptr(lu), lda, float(anorm),
                                    ctypes.byref(out)))
        return out.value

    def dgerfs(self, trans: str, n: int, nrhs: int, a, lda: int, af, ldaf: int, ipiv, b,
               ldb: int, x, ldx: int, ferr=None) -> tuple[list[float], list[float]]:
        """PDGERFS: refines x in place; returns (ferr, berr) per column.  ferr
        keeps the given start values where x_k is zero."""
        fe = (ctypes.c_double * max(1, nrhs))(*(ferr or [0.0] * nrhs))
        be = (ctypes.c_double * max(1, nrhs))()
        _check(self._lib.sla_dgerfs(self._h, trans.encode()[:1], n, nrhs, _ptr(a), lda, _ptr(af),
                                    ldaf, _ptr(ipiv), _ptr(b), ldb, _ptr(x), ldx, fe, be))
        return list(fe[:nrhs]), list(be[:nrhs])

    # ------------------------------------------------------------ batched
    def gesv_batched(self, prec: str, n, nrhs, a_ptrs, lda, ipiv_ptrs, b_ptrs, ldb, info) -> None:
        """Batched small solves: n / nrhs / lda / ldb are sequences of ints;
        a_ptrs / ipiv_ptrs / b_ptrs are device int64 tensors of device
        addresses; info a device int32 tensor (INFO per system)."""
        batch = len(n)
        arr = lambda v: (ctypes.c_int * max(1, batch))(*[int(x) for x in v])  # noqa: E731
        _check(self._lib.sla_gesv_batched(self._h, PRECISIONS[prec], batch, arr(n), arr(nrhs),
                                          _ptr(a_ptrs), arr(lda), _ptr(ipiv_ptrs), _ptr(b_ptrs),
                                          arr(ldb), _ptr(info)))

    def gesv_batched_noalloc(self, prec: str, batch: int, dn, dnrhs, a_ptrs, dlda, ipiv_ptrs,
                             b_ptrs, dldb, info) -> None:
        """sla_gesv_batched_noalloc: sizes as device int32 tensors, nothing
        allocated, work only enqueued on the handle's stream."""
        _check(self._lib.sla_gesv_batched_noalloc(
            self._h, PRECISIONS[prec], batch, _ptr(dn), _ptr(dnrhs), _ptr(a_ptrs), _ptr(dlda),
            _ptr(ipiv_ptrs), _ptr(b_ptrs), _ptr(dldb), _ptr(info)))

    def gesv_batched_tune(self, prec: str, n, nrhs) -> tuple[int, int, float]:
        """Autotune (panel width, workgroups per CU) for these sizes; the
        handle keeps the result.  Returns (w, wgs_per_cu, ms)."""
        batch = len(n)
        arr = lambda v: (ctypes.c_int * max(1, batch))(*[int(x) for x in v])  # noqa: E731
        w, k, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.sla_gesv_batched_tune(self._h, PRECISIONS[prec], batch, arr(n), arr(nrhs),
                                               ctypes.byref(w), ctypes.byref(k), ctypes.byref(ms)))
        return w.value, k.value, ms.value

    # ----------------------------------------------------- learned selector
    def selector_features(self, n: int, a, lda: int):
        """The selector's 12 features of an n x n FP64 device matrix."""
        out = (ctypes.c_float * 12)()
        _check(self._lib.sla_selector_features(self._h, n, _ptr(a), lda, out))
        return list(out)

    def mlp_grad(self, x, y, params, grad, cost=None) -> float:
        """Loss and gradient (into grad) of the selector network; device tensors."""
        loss = ctypes.c_float()
        _check(self._lib.sla_mlp_grad(self._h, x.shape[0], _ptr(x), _ptr(y),
                                      _ptr(cost) if cost is not None else None, _ptr(params),
                                      _ptr(grad), ctypes.byref(loss)))
        return loss.value

    def mlp_train(self, x, y, params, epochs: int, lr: float, cost=None) -> list[float]:
        """Full-batch Adam on params (device, in place); returns the loss history."""
        hist = (ctypes.c_float * max(1, epochs))()
        _check(self._lib.sla_mlp_train(self._h, x.shape[0], _ptr(x), _ptr(y),
                                       _ptr(cost) if cost is not None else None, _ptr(params),
                                       epochs, ctypes.c_float(lr), hist))
        return list(hist[:epochs])

    def mlp_predict(self, x, params, pred=None, prob=None) -> None:
        _check(self._lib.sla_mlp_predict(self._h, x.shape[0], _ptr(x), _ptr(params),
                                         _ptr(pred) if pred is not None else None,
                                         _ptr(prob) if prob is not None else None))

    def selector_set_model(self, params) -> None:
        """params: a host float32 array of SLA_MLP_NPARAM values."""
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        _check(self._lib.sla_selector_set_model(self._h, ctypes.c_void_p(p.ctypes.data)))

    def dxgesv_auto(self, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int, x,
                    ldx: int) -> tuple[int, int, int]:
        """sla_dxgesv with the selector's precision: (choice, iter, info)."""
        c, it, info = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.sla_dxgesv_auto(self._h, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb,
                                         _ptr(x), ldx, ctypes.byref(c), ctypes.byref(it),
                                         ctypes.byref(info)))
        return c.value, it.value, info.value

    # ------------------------------------------------------- mixed precision
    def dxgesv(self, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int, x, ldx: int,
               prec: str = "fp32", refine: str = "ir") -> tuple[int, int]:
        """X = A^{-1} B by a `prec` factorization refined to FP64 backward
        error; returns (iter, info) as LAPACK DSGESV's ITER and INFO."""
        it, info = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.sla_dxgesv(self._h, PRECISIONS[prec], REFINEMENTS[refine], n, nrhs,
                                    _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb, _ptr(x), ldx,
                                    ctypes.byref(it), ctypes.byref(info)))
        return it.value, info.value

    def last_solve_info(self) -> dict:
        out = SolveInfo()
        _check(self._lib.sla_last_solve_info(self._h, ctypes.byref(out)))
        d = out.as_dict()
        d["precision"] = {v: k for k, v in PRECISIONS.items()}[d["precision"]]
        d["refinement"] = {v: k for k, v in REFINEMENTS.items()}[d["refinement"]]
        d["stop_rule"] = STOP_RULES[d["stop_rule"]]
        return d

    def sgetrf(self, m: int, n: int, a, lda: int, ipiv, prec: str = "fp32") -> int:
        """FP32-storage LU with trailing updates in `prec` (fp32, bf16, fp8)."""
        info = ctypes.c_int()
        _check(self._lib.sla_sgetrf(self._h, PRECISIONS[prec], m, n, _ptr(a), lda, _ptr(ipiv),
                                    ctypes.byref(info)))
        return info.value

    def gemm_lowp_sub(self, m: int, n: int, k: int, a, lda: int, b, ldb: int, c, ldc: int,
                      prec: str = "fp32") -> None:
        """C -= A B, FP32 operands, products in `prec`, FP32 accumulation."""
        _check(self._lib.sla_gemm_lowp_sub(self._h, PRECISIONS[prec], m, n, k, _ptr(a), lda,
                                           _ptr(b), ldb, _ptr(c), ldc))

    def sgetrs(self, n: int, nrhs: int, a, lda: int, ipiv, b, ldb: int) -> None:
        """B = A^{-1} B with the FP32 factors of sgetrf."""
        _check(self._lib.sla_sgetrs(self._h, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb))

    def dresidual(self, n: int, nrhs: int, a, lda: int, x, ldx: int, b, ldb: int, r,
                  ldr: int) -> list[float]:
        """R = B - A X in FP64; returns ||R(:, j)||_inf per column."""
        out = (ctypes.c_double * max(1, nrhs))()
        _check(self._lib.sla_dresidual(self._h, n, nrhs, _ptr(a), lda, _ptr(x), ldx, _ptr(b),
                                       ldb, _ptr(r), ldr, out))
        return list(out[:nrhs])
