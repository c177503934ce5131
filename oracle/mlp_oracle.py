"""CPU restatement of the precision selector's network (TEST INFRASTRUCTURE:
only tests use it, as the checker of the HIP forward / backward kernels).

12 -> 64 -> 64 -> 7, tanh hidden layers, softmax output, cost-weighted
cross-entropy averaged over the batch:
    loss = (1/B) sum_s cost[y_s] * -log softmax(z_s)[y_s]
Parameter vector layout: W1[64][12], b1[64], W2[64][64], b2[64], W3[7][64],
b3[7], mu[12], sigma[12] (inputs are (f - mu) / sigma).  Evaluated in
float64 so that it can check the float32 kernels and itself be pinned by
finite differences.
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 12, 64, 7
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)),
         ("b3", (OUT,)), ("mu", (IN,)), ("sd", (IN,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)
NTRAIN = NPARAM - 2 * IN


def unpack(p: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[off:off + n], dtype=np.float64).reshape(shape)
        off += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def forward(p: np.ndarray, X: np.ndarray):
    d = unpack(p)
    x = (np.asarray(X, np.float64) - d["mu"]) / d["sd"]
    h1 = np.tanh(x @ d["W1"].T + d["b1"])
    h2 = np.tanh(h1 @ d["W2"].T + d["b2"])
    z = h2 @ d["W3"].T + d["b3"]
    return x, h1, h2, z


def softmax(z):
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def loss_grad(p: np.ndarray, X: np.ndarray, y: np.ndarray, cost=None):
    """(loss, gradient over the trained parameters: NTRAIN values)."""
    d = unpack(p)
    B = X.shape[0]
    w = np.ones(B) if cost is None else np.asarray(cost, np.float64)[y]
    x, h1, h2, z = forward(p, X)
    pr = softmax(z)
    loss = float(np.sum(-w * np.log(np.maximum(pr[np.arange(B), y], 1e-300))) / B)
    dz = pr.copy()
    dz[np.arange(B), y] -= 1.0
    dz *= (w / B)[:, None]
    g = {"W3": dz.T @ h2, "b3": dz.sum(0)}
    d2 = (dz @ d["W3"]) * (1 - h2 ** 2)
    g["W2"], g["b2"] = d2.T @ h1, d2.sum(0)
    d1 = (d2 @ d["W2"]) * (1 - h1 ** 2)
    g["W1"], g["b1"] = d1.T @ x, d1.sum(0)
    g["mu"], g["sd"] = np.zeros(IN), np.zeros(IN)
    return loss, pack(g)[:NTRAIN]


def loss_only(p, X, y, cost=None) -> float:
    B = X.shape[0]
    w = np.ones(B) if cost is None else np.asarray(cost, np.float64)[y]
    pr = softmax(forward(p, X)[3])
    return float(np.sum(-w * np.log(np.maximum(pr[np.arange(B), y], 1e-300))) / B)
