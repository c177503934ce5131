"""GPU checks of the learned precision selector, through the C-ABI.

* The network's forward and backward kernels against the CPU oracle
  (oracle/mlp_oracle.py, itself pinned by finite differences): FP32
  gradients within 1e-4 of max |g|, probabilities within 1e-5.
* Training is reproducible bit for bit and lowers the loss.
* The 12 features against the CPU restatement (oracle/selector_oracle.py):
  norms and spreads within 1e-5 (log10, FP32 output), sampled features
  within 1e-4.
* The shipped model (scalapack_amd/data/selector_model.npz) on the held-out
  part of the method-generated data set (scalapack_amd/data/
  selector_dataset.npz, held-out systems from (U, V) pools never used in
  training): its choices, feature extraction included, cost less wall time
  than the method without the network (the library's default policy: FP32
  factors, IR continuing as GMRES-IR, FP64 fallback) and less than always
  solving in FP64 -- the gains this test reports and requires.
* sla_dxgesv_auto solves to the testers' scaled residual < 16 with the
  choice the network makes on the same features.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402
import selector_oracle as so  # noqa: E402

DATA = ROOT / "scalapack_amd" / "data"


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


def _params(seed):
    from scalapack_amd.selector import init_params
    rng = np.random.default_rng(seed)
    return init_params(rng.normal(0, 1, 12), rng.uniform(0.5, 2, 12), seed)


@pytest.mark.parametrize("B,cost", [(1, None), (37, [1.0, 2.0, 0.5, 3.0, 1.5, 0.7, 2.5]), (700, None)])
def test_mlp_gradient_matches_oracle(dev, B, cost):
    _note(f"mlp grad B={B}")
    rng = np.random.default_rng(B)
    p = _params(B)
    X = rng.normal(0, 2, (B, 12)).astype(np.float32)
    y = rng.integers(0, 7, B).astype(np.int32)
    dX, dy, dp = (torch.from_numpy(v).cuda() for v in (X, y, p))
    dc = torch.tensor(cost, dtype=torch.float32).cuda() if cost else None
    dg = torch.zeros(mo.NPARAM, dtype=torch.float32, device="cuda")
    loss = dev.mlp_grad(dX, dy, dp, dg, cost=dc)
    lo, go = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), y, cost)
    g = dg.cpu().numpy()[: mo.NTRAIN]
    assert abs(loss - lo) <= 1e-5 * max(1.0, abs(lo))
    assert np.allclose(g, go, rtol=0, atol=1e-4 * np.abs(go).max())


def test_mlp_predict_matches_oracle(dev):
    _note("mlp predict")
    rng = np.random.default_rng(9)
    p = _params(9)
    X = rng.normal(0, 2, (300, 12)).astype(np.float32)
    dX, dp = torch.from_numpy(X).cuda(), torch.from_numpy(p).cuda()
    pred = torch.zeros(300, dtype=torch.int32, device="cuda")
    prob = torch.zeros((300, 7), dtype=torch.float32, device="cuda")
    dev.mlp_predict(dX, dp, pred=pred, prob=prob)
    po = mo.softmax(mo.forward(p.astype(np.float64), X.astype(np.float64))[3])
    assert np.allclose(prob.cpu().numpy(), po, rtol=0, atol=1e-5)
    clear = np.sort(po, axis=1)[:, -1] - np.sort(po, axis=1)[:, -2] > 1e-4
    assert np.array_equal(pred.cpu().numpy()[clear], po.argmax(1)[clear])


def test_mlp_training_is_reproducible_and_learns(dev):
    _note("mlp train")
    rng = np.random.default_rng(10)
    X = rng.normal(0, 1, (500, 12)).astype(np.float32)
    y = (X[:, 0] > 0).astype(np.int32) + 2 * (X[:, 3] > 0.5).astype(np.int32)  # learnable
    p0 = _params(11)
    out, hists = [], []
    for _ in range(2):
        dp = torch.from_numpy(p0.copy()).cuda()
        hists.append(dev.mlp_train(torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda(), dp, 300, 1e-2))
        out.append(dp.cpu().numpy())
    assert np.array_equal(out[0], out[1]) and hists[0] == hists[1]
    assert hists[0][-1] < 0.3 * hists[0][0]
    pred = torch.zeros(500, dtype=torch.int32, device="cuda")
    dev.mlp_predict(torch.from_numpy(X).cuda(), torch.from_numpy(out[0]).cuda(), pred=pred)
    assert (pred.cpu().numpy() == y).mean() > 0.95


@pytest.mark.parametrize("n,kappa,spread", [(100, 1e3, 1.0), (777, 1e8, 1e4), (1500, 1e2, 1.0)])
def test_selector_features_match_oracle(dev, n, kappa, spread):
    _note(f"features n={n}")
    rng = np.random.default_rng(n)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / (n - 1))
    a = (spread ** rng.uniform(0, 1, n))[:, None] * ((U * s) @ V.T)
    dA = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    f = np.array(dev.selector_features(n, dA, n))
    fo = so.features(a)
    assert np.allclose(f[:7], fo[:7], rtol=0, atol=1e-5), (f, fo)
    assert np.allclose(f[7:], fo[7:], rtol=0, atol=1e-4), (f, fo)


def _model():
    from scalapack_amd.selector import load_model
    mp, dp = DATA / "selector_model.npz", DATA / "selector_dataset.npz"
    if not (mp.exists() and dp.exists()):
        pytest.skip("the shipped selector model and data set are not built yet")
    return load_model(mp), dict(np.load(dp))


def test_shipped_model_gains_over_fixed_policy(dev):
    from scalapack_amd.selector import evaluate
    p, d = _model()
    ntr = int(d["n_train"])
    X = d["features"][ntr:].astype(np.float32)
    _note(f"selector held-out evaluation ({X.shape[0]} samples)")
    pred = torch.zeros(X.shape[0], dtype=torch.int32, device="cuda")
    dev.mlp_predict(torch.from_numpy(X).cuda(), torch.from_numpy(p).cuda(), pred=pred)
    pr = pred.cpu().numpy()
    ev = evaluate(d, pr)
    po = mo.softmax(mo.forward(p.astype(np.float64), X.astype(np.float64))[3]).argmax(1)
    assert (pr == po).mean() > 0.99          # the kernels' choice is the oracle's
    _note(f"  {ev}")
    assert ev["accuracy"] > 0.5
    # the network (features included) against the library's default policy
    # (FP32 factors, IR continuing as GMRES-IR, FP64 fallback) and against
    # always solving in FP64, on the same held-out systems
    assert ev["gain_vs_default"] > 1.0
    assert ev["gain_vs_fp64"] > 1.0


def test_dxgesv_auto_solves_with_the_network_choice(dev):
    p, _ = _model()
    dev.selector_set_model(p)
    for n, kappa in ((512, 1e2), (1024, 1e10)):
        _note(f"dxgesv_auto n={n} kappa={kappa:g}")
        rng = np.random.default_rng(int(kappa) % 1000 + n)
        U, _ = np.linalg.qr(rng.standard_normal((n, n)))
        V, _ = np.linalg.qr(rng.standard_normal((n, n)))
        a = (U * kappa ** (-np.arange(n) / (n - 1))) @ V.T
        x0 = rng.standard_normal(n)
        b = a @ x0
        AT = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
        f = dev.selector_features(n, AT, n)
        c_expect = torch.zeros(1, dtype=torch.int32, device="cuda")
        dev.mlp_predict(torch.tensor([f], dtype=torch.float32).cuda(), torch.from_numpy(p).cuda(),
                        pred=c_expect)
        A = AT.clone()
        db = torch.from_numpy(b).cuda()
        X = torch.zeros(n, dtype=torch.float64, device="cuda")
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        choice, it, info = dev.dxgesv_auto(n, 1, A, n, ipiv, db, n, X, n)
        assert info == 0 and choice == int(c_expect.item())
        x = X.cpu().numpy()
        r = np.abs(b - a @ x).max() / (np.abs(a).sum(1).max() * np.abs(x).max() * n * 2.0 ** -53)
        assert r < 16.0, (choice, it, r)
