"""GPU tests of the mixed-precision path, through the C-ABI.

* The low-precision trailing-update kernels (FP32, BF16, FP8-E4M3 products,
  FP32 accumulation) against a float64 product of the same rounded operands
  (torch's bfloat16 / float8_e4m3fn rounding, the FP8 per-block scale
  max|block| / 448 applied as in the kernels); tolerance: FP32 accumulation,
  2 k 2^-24 (|A||B|) + 2^-24 |C|.
* The FP32-storage factorization: P A = L U elementwise within a multiple of
  |L||U| set by the update precision (FP32 2 n 2^-24, BF16 2^-5, FP8 0.5:
  operands rounded to 8 and 4 significant bits), |L| <= 1, pivots a valid
  interchange sequence.
* The refined solve: the scaled residual ||b - A x|| / (||A|| ||x|| N eps) < 16
  of the LU testers (TESTING/traditional/LIN/pdlaschk.f:296) for every
  result; the forward error against the oracle's FP64 LU solution is
  reported and bounded by 100x the oracle's own forward error (x_true
  known); the DSGESV contract (A untouched unless ITER < 0, B untouched,
  ITER codes -1 / -2 / -3, INFO of the FP64 fallback).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = 2.0 ** -53


def _note(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from scalapack_amd.device import Handle
    h = Handle()
    yield h
    h.close()


def to_dev(a: np.ndarray, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).T))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def from_dev(t) -> np.ndarray:
    return np.asfortranarray(t.cpu().numpy().T)


def svd_matrix(n: int, kappa: float, seed: int) -> np.ndarray:
    """A = U diag(sigma) V^T, U and V Q factors of seeded Gaussians, sigma
    geometric from 1 to 1/kappa."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / (n - 1))
    return np.asfortranarray((u * s) @ v.T)


def scaled_residual(a, x, b) -> float:
    r = b - a @ x
    an = np.abs(a).sum(axis=1).max()
    n = a.shape[0]
    return float((np.abs(r).max(axis=0) / (an * np.abs(x).max(axis=0) * n * EPS)).max())


def quantize(a32: np.ndarray, prec: str) -> tuple[np.ndarray, float]:
    """Operand as the kernels round it (float64 values) and its scale."""
    t = torch.from_numpy(np.ascontiguousarray(a32))
    if prec == "fp32":
        return a32.astype(np.float64), 1.0
    if prec == "bf16":
        return t.to(torch.bfloat16).to(torch.float64).numpy(), 1.0
    amax = np.float32(np.abs(a32).max())
    scale = np.float32(amax / np.float32(448.0)) if amax > 0 else np.float32(1.0)
    inv = np.float32(np.float32(1.0) / scale)
    q = torch.from_numpy(np.ascontiguousarray((a32 * inv).astype(np.float32)))
    q = q.clamp(-448.0, 448.0).to(torch.float8_e4m3fn).to(torch.float64).numpy()
    return q, float(scale)


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8"])
@pytest.mark.parametrize("m,n,k", [(128, 128, 128), (256, 384, 256), (1024, 1024, 512),
                                   (300, 200, 77), (130, 260, 32)])
def test_gemm_lowp_parity(dev, prec, m, n, k):
    _note(f"gemm_lowp {prec} m={m} n={n} k={k}")
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(np.float32))
    B = np.asfortranarray(rng.uniform(-2, 2, (k, n)).astype(np.float32))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    dev.gemm_lowp_sub(m, n, k, dA, m, dB, k, dC, m, prec=prec)
    torch.cuda.synchronize()
    got = from_dev(dC).astype(np.float64)
    qa, sa = quantize(A, prec)
    qb, sb = quantize(B, prec)
    prod = (qa @ qb) * (sa * sb)
    ref = C.astype(np.float64) - prod
    bound = 2.0 * k * 2.0 ** -24 * (np.abs(qa) @ np.abs(qb)) * (sa * sb) + 2.0 ** -23 * np.abs(C)
    err = np.abs(got - ref)
    assert np.all(err <= bound + 1e-30), f"max err {err.max()} vs bound {bound.max()}"


SPLITS = {"bf16x3": (torch.bfloat16, 2), "bf16x6": (torch.bfloat16, 3),
          "fp16x3": (torch.float16, 2)}
# accuracy of the split product against the exact product of the FP32
# operands, as a fraction of |A||B|: dropped slice products plus the slices'
# representation error (FP32 accumulation adds its own 2 k 2^-24)
SPLIT_REL = {"bf16x3": 2.0 ** -15, "bf16x6": 2.0 ** -22, "fp16x3": 2.0 ** -19}


def split_slices(a32: np.ndarray, prec: str) -> tuple[list[np.ndarray], int]:
    """FP32 operand as the split kernels write it: (FP16: scaled by 2^e so
    that max|a| 2^e lies in [2^14, 2^15)) slices a_i = r16(a - a_0 - ... -
    a_{i-1}), round to nearest even by torch's conversions; returns the
    slices as float64 arrays and e."""
    dt, S = SPLITS[prec]
    e = 0
    if dt == torch.float16:
        amax = float(np.abs(a32).max())
        if amax > 0:
            e = 15 - int(np.frexp(np.float32(amax))[1])
    x = torch.from_numpy(np.ascontiguousarray(np.ldexp(a32, e).astype(np.float32)))
    out = []
    for _ in range(S):
        sl = x.to(dt)
        out.append(sl.to(torch.float64).numpy())
        x = x - sl.to(torch.float32)
    return out, e


@pytest.mark.parametrize("prec", ["bf16x3", "bf16x6", "fp16x3"])
@pytest.mark.parametrize("m,n,k", [(128, 128, 128), (256, 384, 512), (1024, 1024, 512),
                                   (300, 200, 77), (130, 260, 32), (520, 390, 64),
                                   (768, 640, 512)])
def test_gemm_split_parity(dev, prec, m, n, k):
    """Split 16-bit products: against the float64 sum of the slice products
    the kernel forms (FP32 accumulation bound), and against the exact product
    of the FP32 operands (the split's accuracy claim)."""
    _note(f"gemm_lowp {prec} m={m} n={n} k={k}")
    rng = np.random.default_rng(m + 5 * n + 11 * k)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, k)).astype(np.float32))
    B = np.asfortranarray((rng.uniform(-2, 2, (k, n)) * 10.0 ** rng.uniform(-3, 0, (k, 1)))
                          .astype(np.float32))
    C = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    dev.gemm_lowp_sub(m, n, k, dA, m, dB, k, dC, m, prec=prec)
    torch.cuda.synchronize()
    got = from_dev(dC).astype(np.float64)
    sa, ea = split_slices(A, prec)
    sb, eb = split_slices(B, prec)
    S = len(sa)
    prod = sum(sa[i] @ sb[j] for i in range(S) for j in range(S) if i + j < S)
    prod = np.ldexp(prod, -(ea + eb))
    mag = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    ref = C.astype(np.float64) - prod
    err = np.abs(got - ref)
    bound = 3.0 * S * k * 2.0 ** -24 * mag + 2.0 ** -23 * np.abs(C)
    assert np.all(err <= bound + 1e-30), f"max err {err.max()} vs bound {bound.max()}"
    exact = C.astype(np.float64) - A.astype(np.float64) @ B.astype(np.float64)
    rel = (np.abs(got - exact) - 2.0 ** -23 * np.abs(C) - 2 * k * 2.0 ** -24 * mag) / mag
    _note(f"  max |split - exact| / |A||B| beyond FP32 accumulation: {rel.max():.3e}")
    assert rel.max() <= SPLIT_REL[prec]


# --------------------------------------------------------- factorization
@pytest.mark.parametrize("pivoting", [0, 1])
@pytest.mark.parametrize("prec", ["fp32", "bf16", "fp8", "bf16x3", "bf16x6", "fp16x3"])
@pytest.mark.parametrize("n,nb", [(1024, 256), (1000, 256), (2048, 512), (1536, 128)])
def test_sgetrf_backward_error(dev, prec, n, nb, pivoting):
    _note(f"sgetrf {prec} n={n} nb={nb} pivoting={pivoting}")
    tol = {"fp32": 2 * n * 2.0 ** -24, "bf16": 2.0 ** -5, "fp8": 0.5, "bf16x3": 2.0 ** -11,
           "bf16x6": 2 * n * 2.0 ** -24, "fp16x3": 2.0 ** -12}[prec]
    rng = np.random.default_rng(n + len(prec))
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
    dA = to_dev(A)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    nb0, piv0 = dev.get("nb"), dev.get("pivoting")
    dev.set("nb", nb)
    dev.set("pivoting", pivoting)
    try:
        info = dev.sgetrf(n, n, dA, n, ipiv, prec=prec)
    finally:
        dev.set("nb", nb0)
        dev.set("pivoting", piv0)
    torch.cuda.synchronize()
    assert info == 0
    lu = from_dev(dA).astype(np.float64)
    piv = ipiv.cpu().numpy()
    assert np.all(piv >= np.arange(1, n + 1)) and np.all(piv <= n)
    perm = np.arange(n)
    for i, p in enumerate(piv):
        perm[[i, p - 1]] = perm[[p - 1, i]]
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    err = (np.abs(A.astype(np.float64)[perm] - L @ U) / (np.abs(L) @ np.abs(U))).max()
    _note(f"  max |PA - LU| / (|L||U|) = {err:.3e} (bound {tol:.1e})")
    assert err <= tol
    lmax = np.abs(L).max()
    _note(f"  max |L| = {lmax:.3f}")
    if pivoting == 0:
        assert lmax <= 1.0 + 1e-6  # partial pivoting keeps |l| <= 1
    else:
        assert lmax <= 8.0  # tournament pivoting: bounded multipliers in practice


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6", "fp8"])
@pytest.mark.parametrize("m,n,nb", [(2560, 2560, 512), (1700, 1100, 128), (900, 1400, 128)])
def test_sgetrf_lookahead_matches_serial_order(dev, prec, m, n, nb):
    """The FP32-storage factorization with the look-ahead (panel stream
    beside the update stream) against the serial launch order (lookahead =
    0): the same arithmetic -- tournament leaves, packed low-precision
    updates, every row interchange on the update stream -- in another order
    of launches, so IPIV and the factors are bit-identical."""
    _note(f"sgetrf lookahead {prec} m={m} n={n} nb={nb}")
    rng = np.random.default_rng(m + 3 * n + nb)
    A = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(np.float32))
    mn = min(m, n)
    out = {}
    nb0, la0 = dev.get("nb"), dev.get("lookahead")
    dev.set("nb", nb)
    try:
        for la in (0, 1):
            dev.set("lookahead", la)
            dA = to_dev(A)
            ipiv = torch.zeros(mn, dtype=torch.int32, device="cuda")
            assert dev.sgetrf(m, n, dA, m, ipiv, prec=prec) == 0
            torch.cuda.synchronize()
            out[la] = (ipiv.cpu().numpy(), from_dev(dA))
    finally:
        dev.set("lookahead", la0)
        dev.set("nb", nb0)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


# ----------------------------------------------------------- refined solve
def _solve(dev, A, B, prec, refine):
    n, nrhs = B.shape
    dA, dB = to_dev(A), to_dev(B)
    dA0, dB0 = dA.clone(), dB.clone()
    dX = torch.zeros_like(dB)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    it, info = dev.dxgesv(n, nrhs, dA, n, ipiv, dB, n, dX, n, prec=prec, refine=refine)
    torch.cuda.synchronize()
    st = dev.last_solve_info()
    assert torch.equal(dB, dB0), "B was modified"
    if it >= 0:
        assert torch.equal(dA, dA0), "A was modified although refinement succeeded"
    return from_dev(dX), it, info, st, dA, ipiv


def _oracle_x(oracle, A, B, nb=64):
    lu = A.copy(order="F")
    piv, info = oracle.pdgetrf(lu, nb)
    assert info == 0
    return oracle.pdgetrs(lu, piv, B.copy(order="F"), nb)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("n,nrhs", [(2048, 1), (1500, 3), (1024, 16), (4096, 1), (1001, 1),
                                    (777, 3)])
def test_dxgesv_uniform(dev, oracle, prec, n, nrhs):
    """(1001, 1) and (777, 3): n nrhs odd, so 12 n nrhs bytes of R and W are
    not a multiple of 8 -- the norm words after them must still be aligned
    for their 64-bit atomics (the refinement buffer's layout)."""
    _note(f"dxgesv {prec} IR uniform n={n} nrhs={nrhs}")
    rng = np.random.default_rng(n + nrhs)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    xt = np.asfortranarray(rng.standard_normal((n, nrhs)))
    B = np.asfortranarray(A @ xt)
    X, it, info, st, _, _ = _solve(dev, A, B, prec, "ir")
    assert info == 0
    s = scaled_residual(A, X, B)
    x64 = _oracle_x(oracle, A, B)
    fe = np.abs(X - xt).max() / np.abs(xt).max()
    fe64 = np.abs(x64 - xt).max() / np.abs(xt).max()
    _note(f"  iter={it} {st} sresid={s:.3e} fwd={fe:.3e} fwd(FP64 LU)={fe64:.3e}")
    assert s < 16.0
    if it >= 0:
        assert abs(st["sresid"] - s) <= 0.5 * s + 1.0  # the library's own measure agrees
    assert fe <= 100 * fe64 + 1e-14


@pytest.mark.parametrize("prec,refine,kappa", [("fp32", "gmres", 1e8), ("fp32", "ir", 1e4),
                                               ("bf16", "gmres", 1e3), ("fp32", "gmres", 1e6)])
def test_dxgesv_ill_conditioned(dev, oracle, prec, refine, kappa):
    n = 1024
    _note(f"dxgesv {prec} {refine} kappa={kappa:g} n={n}")
    A = svd_matrix(n, kappa, 7)
    rng = np.random.default_rng(8)
    xt = np.asfortranarray(rng.standard_normal((n, 1)))
    B = np.asfortranarray(A @ xt)
    X, it, info, st, _, _ = _solve(dev, A, B, prec, refine)
    assert info == 0
    s = scaled_residual(A, X, B)
    x64 = _oracle_x(oracle, A, B)
    fe = np.abs(X - xt).max() / np.abs(xt).max()
    fe64 = np.abs(x64 - xt).max() / np.abs(xt).max()
    _note(f"  iter={it} {st} sresid={s:.3e} fwd={fe:.3e} fwd(FP64 LU)={fe64:.3e}")
    assert s < 16.0
    assert it >= 0, "refinement fell back to FP64"
    assert fe <= 100 * fe64 + 1e-14


@pytest.mark.parametrize("prec,refine,kappa", [("bf16", "gmres", 1e2), ("bf16", "gmres", 1e3),
                                               ("bf16", "ir", 1e1), ("fp8", "gmres", 1e1),
                                               ("bf16x3", "ir", 1e3), ("fp16x3", "ir", 1e4),
                                               ("bf16x6", "ir", 1e4), ("bf16x6", "gmres", 1e6)])
def test_dxgesv_low_precision_converges(dev, oracle, prec, refine, kappa):
    """Where kappa times the factors' relative error stays well below 1
    (BF16 updates: 8 significant bits; FP8-E4M3 with per-block scales: 4;
    split products: ~2^-16, ~2^-21, ~2^-24) the refinement itself must reach
    FP64 backward error: no FP64 fallback (ITER >= 0), DSGESV's convergence
    test or an accepted stall, and a forward error within 100x that of the
    oracle's FP64 LU.  (FP8 + GMRES-IR at kappa = 1e2 does not: one 50-step
    correction only takes the backward error from 1e-2 to 7e-3, and the FP64
    fallback answers -- covered by the next test.)"""
    n = 1024
    _note(f"dxgesv {prec} {refine} kappa={kappa:g} n={n} (must converge)")
    A = svd_matrix(n, kappa, 13)
    rng = np.random.default_rng(14)
    xt = np.asfortranarray(rng.standard_normal((n, 1)))
    B = np.asfortranarray(A @ xt)
    X, it, info, st, _, _ = _solve(dev, A, B, prec, refine)
    assert info == 0
    s = scaled_residual(A, X, B)
    x64 = _oracle_x(oracle, A, B)
    fe = np.abs(X - xt).max() / np.abs(xt).max()
    fe64 = np.abs(x64 - xt).max() / np.abs(xt).max()
    _note(f"  iter={it} {st} sresid={s:.3e} fwd={fe:.3e} fwd(FP64 LU)={fe64:.3e}")
    assert it >= 0 and st["fell_back"] == 0, "refinement fell back to FP64"
    assert st["stop_rule"] in ("converged", "stall_accepted")
    assert s < 16.0
    assert fe <= 100 * fe64 + 1e-14


def test_dxgesv_stall_acceptance_is_opt_out(dev):
    """ir_accept_stall = 0: a stalled refinement is never returned as a
    result (the FP64 fallback answers instead); with the default it may be,
    and stop_rule says which rule ended the refinement."""
    n = 1024
    A = svd_matrix(n, 1e4, 21)
    B = np.asfortranarray(A @ np.ones((n, 1)))
    dev.set("ir_accept_stall", 0)
    try:
        X, it, info, st, _, _ = _solve(dev, A, B, "bf16", "ir")
    finally:
        dev.set("ir_accept_stall", 1)
    _note(f"  strict: iter={it} stop={st['stop_rule']}")
    assert st["stop_rule"] in ("converged", "fallback")
    assert (it >= 0) == (st["stop_rule"] == "converged")
    assert scaled_residual(A, X, B) < 16.0


@pytest.mark.parametrize("prec,kappa", [("fp8", 1e6), ("fp8", 1e2), ("bf16", 1e4)])
def test_dxgesv_low_precision_falls_back_or_converges(dev, oracle, prec, kappa):
    """kappa times the unit roundoff of the factors above 1 (FP8: 2^-4,
    BF16: 2^-8) may stall GMRES-IR; either way the answer meets the FP64
    criterion, from the refinement or from the FP64 fallback."""
    n = 1024
    _note(f"dxgesv {prec} gmres kappa={kappa:g} n={n}")
    A = svd_matrix(n, kappa, 11)
    xt = np.ones((n, 1))
    B = np.asfortranarray(A @ xt)
    X, it, info, st, dA, ipiv = _solve(dev, A, B, prec, "gmres")
    assert info == 0
    s = scaled_residual(A, X, B)
    _note(f"  iter={it} {st} sresid={s:.3e}")
    assert s < 16.0


def test_dxgesv_overflow_falls_back(dev, oracle):
    n = 512
    _note("dxgesv overflow -> FP64 fallback")
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)) + n * np.eye(n))
    A[5, 5] = 1e39
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 1)))
    X, it, info, st, dA, ipiv = _solve(dev, A, B, "fp32", "ir")
    assert (it, info, st["fell_back"]) == (-2, 0, 1)
    assert scaled_residual(A, X, B) < 16.0
    lu = A.copy(order="F")
    piv, _ = oracle.pdgetrf(lu, 64)
    assert np.array_equal(ipiv.cpu().numpy(), piv[:n])


def test_dxgesv_zero_pivot(dev, oracle):
    n = 256
    _note("dxgesv exactly singular -> ITER -3, INFO of FP64 LU")
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    A[:, 100] = 0.0
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 1)))
    lu = A.copy(order="F")
    _, info_ref = oracle.pdgetrf(lu, 64)
    _, it, info, st, _, _ = _solve(dev, A, B, "fp32", "ir")
    assert it == -3 and info == info_ref == 101


def test_dxgesv_fp64_request(dev, oracle):
    n = 384
    _note("dxgesv prec=fp64 -> plain FP64 LU, ITER -1")
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.uniform(-1, 1, (n, n)))
    B = np.asfortranarray(rng.uniform(-1, 1, (n, 2)))
    X, it, info, st, _, _ = _solve(dev, A, B, "fp64", "ir")
    assert (it, info) == (-1, 0)
    x64 = _oracle_x(oracle, A, B)
    assert np.allclose(X, x64, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("pivoting", [0, 1])
def test_kappa_1e8_pivoting_does_not_decide_convergence(dev, oracle, pivoting):
    """N = 2048, kappa = 1e8, FP32 factors + GMRES-IR: partial pivoting
    (pivoting = 0, |l| <= 1) and the tournament leaf (pivoting = 1) end the
    same way -- the refinement does not reach FP64 backward error and the
    FP64 fallback answers (ITER < 0, scaled residual < 16, forward error
    within 100x the oracle's FP64 LU).  At this kappa it is the FP32
    factors' accuracy, not the pivoting's growth, that stops the refinement
    (at N = 1024 both converge: test_dxgesv_ill_conditioned)."""
    n = 2048
    _note(f"dxgesv fp32 gmres kappa=1e8 n={n} pivoting={pivoting}")
    A = svd_matrix(n, 1e8, 7)
    rng = np.random.default_rng(8)
    xt = np.asfortranarray(rng.standard_normal((n, 1)))
    B = np.asfortranarray(A @ xt)
    piv0 = dev.get("pivoting")
    dev.set("pivoting", pivoting)
    try:
        X, it, info, st, _, _ = _solve(dev, A, B, "fp32", "gmres")
    finally:
        dev.set("pivoting", piv0)
    s = scaled_residual(A, X, B)
    x64 = _oracle_x(oracle, A, B)
    fe = np.abs(X - xt).max() / np.abs(xt).max()
    fe64 = np.abs(x64 - xt).max() / np.abs(xt).max()
    _note(f"  iter={it} stop={st.get('stop_rule')} sresid={s:.3e} fwd={fe:.3e} fwd64={fe64:.3e}")
    assert info == 0 and it < 0
    assert s < 16.0
    assert fe <= 100 * fe64 + 1e-14


@pytest.mark.parametrize("pivoting", [0, 1])
def test_dxgesv_large_fp32_ir(dev, pivoting):
    """N = 8192, FP32 factors, classical refinement; checked on the GPU in
    FP64 (no oracle at this size): scaled residual < 16 and the solution of
    b = A ones within the conditioning of A."""
    n = 8192
    _note(f"dxgesv fp32 IR uniform n={n} pivoting={pivoting}")
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    AT = torch.rand((n, n), generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    B = (AT.T @ torch.ones(n, 1, device="cuda", dtype=torch.float64)).T.contiguous()
    X = torch.zeros_like(B)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    piv0 = dev.get("pivoting")
    dev.set("pivoting", pivoting)
    try:
        it, info = dev.dxgesv(n, 1, AT, n, ipiv, B, n, X, n, prec="fp32", refine="ir")
    finally:
        dev.set("pivoting", piv0)
    torch.cuda.synchronize()
    st = dev.last_solve_info()
    R = B.T - AT.T @ X.T
    an = AT.abs().sum(dim=0).max().item()
    s = (R.abs().max() / (an * X.abs().max() * n * EPS)).item()
    _note(f"  iter={it} history={st['history']} sresid={s:.3e}")
    assert info == 0 and it >= 0
    assert s < 16.0


@pytest.mark.parametrize("blockinv", [0, 1])
@pytest.mark.parametrize("n,nrhs", [(1, 1), (100, 3), (128, 1), (129, 2), (1000, 16),
                                    (4100, 1), (4100, 5)])
def test_sgetrs_matches_float64_solve_with_the_same_factors(dev, n, nrhs, blockinv):
    """sla_sgetrs (FP32 arithmetic) against P^T-applied float64 triangular
    solves with the same FP32 factors of a diagonally dominant A: relative
    error within 64 n 2^-24 (the FP32 solves' rounding; the factors are
    well conditioned).  Both solvers: substitution (blockinv 0) and the
    inverses of the 128 x 128 diagonal blocks (blockinv 1, block edges at
    128 and 129, several block rows at 4100)."""
    _note(f"sgetrs n={n} nrhs={nrhs} blockinv={blockinv}")
    rng = np.random.default_rng(n + 17 * nrhs)
    A = rng.uniform(-1, 1, (n, n)) + n * np.eye(n)
    B = rng.uniform(-1, 1, (n, nrhs))
    dA = to_dev(A, torch.float32)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    assert dev.sgetrf(n, n, dA, n, ipiv, prec="fp32") == 0
    dB = to_dev(B, torch.float32)
    b0 = dev.get("trsv_blockinv")
    dev.set("trsv_blockinv", blockinv)
    try:
        dev.sgetrs(n, nrhs, dA, n, ipiv, dB, n)
    finally:
        dev.set("trsv_blockinv", b0)
    torch.cuda.synchronize()
    LU = from_dev(dA).astype(np.float64)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    pb = B.copy()
    for i, p in enumerate(ipiv.cpu().numpy()):  # PDLAPIV: interchanges in order
        pb[[i, p - 1]] = pb[[p - 1, i]]
    import scipy.linalg as sl
    x_ref = sl.solve_triangular(U, sl.solve_triangular(L, pb, lower=True, unit_diagonal=True))
    X = from_dev(dB).astype(np.float64)
    err = np.abs(X - x_ref).max() / np.abs(x_ref).max()
    _note(f"  max rel error {err:.3e}")
    assert err <= 64 * n * 2.0 ** -24
