"""CPU restatement of the precision selector's 12 features (TEST
INFRASTRUCTURE: the checker of the GPU feature extraction).

The sampled features factor four principal k x k submatrices (k = min(n,
128); one when n <= 128) whose index sets come from a partial Fisher-Yates
shuffle driven by the 64-bit LCG x <- 6364136223846793005 x +
1442695040888963407 seeded with 0x9E3779B97F4A7C15 (t + 1) + n, index i
swapped with i + (x >> 33) mod (n - i); LU with partial pivoting (LAPACK
getrf through scipy) gives the growth max |L\\U| / max |A_s| and the diagonal
ratio max |u_ii| / min |u_ii|; the PDGECON-style estimate is ||A_s||_1 *
2 ||A_s^{-1} x_alt||_1 / (3k) with x_alt the alternating-sign vector of
SRC/pdlacon.f (label 120).  The last feature is the stable-rank fraction
log10(||A||_F^2 / (n lambda)), lambda the Rayleigh estimate of ||A||_2^2
from two power steps on A^T A started at an integer-hash vector (z1 = A^T A v
scaled to max |z1| = 1, z2 = A^T A z1, lambda = z1.z2 / z1.z1).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

M64 = (1 << 64) - 1


def sample_indices(n: int) -> list[np.ndarray]:
    k = min(n, 128)
    S = 4 if n > 128 else 1
    out = []
    for t in range(S):
        perm = list(range(n))
        st = (0x9E3779B97F4A7C15 * (t + 1) + n) & M64
        for i in range(k):
            st = (st * 6364136223846793005 + 1442695040888963407) & M64
            j = i + (st >> 33) % (n - i)
            perm[i], perm[j] = perm[j], perm[i]
        out.append(np.sort(np.array(perm[:k])))
    return out


def hash_vector(n: int) -> np.ndarray:
    """The power steps' start: x_i in [-1, 1) from a 32-bit integer hash of i."""
    v = (np.arange(n, dtype=np.uint64) * 2654435761 + 0x9E3779B9) & 0xFFFFFFFF
    v ^= v >> 16
    v = (v * 0x7feb352d) & 0xFFFFFFFF
    v ^= v >> 15
    v = (v * 0x846ca68b) & 0xFFFFFFFF
    v ^= v >> 16
    return v.astype(np.float64) / 2147483648.0 - 1.0


def power_lambda(a: np.ndarray) -> float:
    z1 = a.T @ (a @ hash_vector(a.shape[0]))
    m = np.abs(z1).max()
    if not (m > 0 and np.isfinite(m)):
        return 0.0
    z1 = z1 / m
    z2 = a.T @ (a @ z1)
    d1 = z1 @ z1
    return float(z1 @ z2 / d1) if d1 > 0 else 0.0


def _lg(x: float) -> float:
    return float(np.log10(min(max(x, 1e-300), 1e300)))


def features(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float64)
    n = a.shape[0]
    f = np.zeros(12)
    f[0] = np.log10(n)
    rows, cols = np.abs(a).sum(1), np.abs(a).sum(0)
    f[1], f[2], f[3] = _lg(cols.max()), _lg(rows.max()), _lg(np.abs(a).max())
    f[4] = _lg(np.linalg.norm(a) / np.sqrt(n))
    f[5] = _lg(rows.max()) - _lg(rows.min())
    f[6] = _lg(cols.max()) - _lg(cols.min())
    g, d, c = [], [], []
    for ix in sample_indices(n):
        s = a[np.ix_(ix, ix)]
        k = s.shape[0]
        lu, piv = sla.lu_factor(s, check_finite=False)
        diag = np.abs(np.diag(lu))
        g.append(_lg(np.abs(lu).max()) - _lg(np.abs(s).max()))
        d.append(min(20.0, _lg(diag.max()) - _lg(diag.min())))
        kk = np.arange(1, k + 1)
        alt = np.where(kk % 2 == 0, -1.0, 1.0) * (1.0 + (kk - 1) / max(k - 1, 1))
        if diag.min() == 0.0:
            c.append(20.0)
        else:
            x = sla.lu_solve((lu, piv), alt)
            c.append(min(20.0, _lg(np.abs(s).sum(0).max() * 2.0 * np.abs(x).sum() / (3.0 * k))))
    f[7], f[8], f[9], f[10] = np.mean(g), max(d), np.mean(c), max(c)
    lam = power_lambda(a)
    fro2 = np.linalg.norm(a) ** 2
    f[11] = min(20.0, max(-20.0, _lg(fro2 / (n * lam)))) if lam > 0 else -20.0
    return f
