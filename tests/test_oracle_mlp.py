"""The CPU oracle of the selector network (oracle/mlp_oracle.py) pinned by
the mathematics: its analytic gradient against central finite differences
in float64 (a dropped tanh derivative, a transposed weight or a missing
class weight fails), the softmax probabilities summing to one, and the
loss of a zero network being log 7 (seven classes)."""
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mo.NPARAM)
    p[-24:-12] = rng.normal(0, 1, 12)          # mu
    p[-12:] = rng.uniform(0.5, 2.0, 12)        # sigma
    return p


@pytest.mark.parametrize("cost", [None, [1.0, 2.0, 0.5, 3.0, 1.5, 0.7, 2.5]])
def test_gradient_matches_finite_differences(cost):
    rng = np.random.default_rng(3)
    p = _params(4)
    X = rng.normal(0, 2, (9, 12))
    y = rng.integers(0, 7, 9)
    loss, g = mo.loss_grad(p, X, y, cost)
    assert np.isclose(loss, mo.loss_only(p, X, y, cost), rtol=1e-14, atol=0)
    idx = rng.choice(mo.NTRAIN, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss_only(p + e, X, y, cost) - mo.loss_only(p - e, X, y, cost)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_softmax_and_zero_network():
    p = _params(5)
    X = np.random.default_rng(6).normal(size=(5, 12))
    pr = mo.softmax(mo.forward(p, X)[3])
    assert np.allclose(pr.sum(1), 1.0, rtol=0, atol=1e-15)
    z = np.zeros(mo.NPARAM)
    z[-12:] = 1.0
    loss, g = mo.loss_grad(z, X, np.array([0, 1, 2, 3, 0]))
    assert np.isclose(loss, np.log(7.0), rtol=1e-15, atol=0)
    assert mo.NPARAM == 5471 and mo.NTRAIN == 5447
