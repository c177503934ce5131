"""GPU parity of the batched small solves (one persistent launch), through
the C-ABI.

Each system is checked against the oracle's PDGESV on a 1 x 1 grid with NB >=
N (the unblocked PDGETF2 sequence): IPIV equal except at near-ties (relative
gap of the two largest pivot candidates below 1e-10, recorded by the
oracle), X within 1e-10 of max |x| for these well-conditioned matrices, and
the tester's scaled residual ||A x - b|| / (||A|| ||x|| n eps) < 16
(TESTING/traditional/LIN/pdlaschk.f:296) for FP64; FP32 systems against the
same criterion with eps = 2^-24.  Zero pivots set INFO to the column and
leave B untouched (SRC/pdgesv.f:231).  Graph replays must equal the eager
launch bit for bit (same kernel, same inputs, fixed operation order).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


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


class Batch:
    """Systems packed in one device buffer each for A, B, IPIV, with the
    pointer arrays the C-ABI takes."""

    def __init__(self, mats, rhs, dtype=torch.float64, pad=0):
        self.n = [a.shape[0] for a in mats]
        self.nrhs = [b.shape[1] for b in rhs]
        self.lda = [max(1, n + pad) for n in self.n]
        self.ldb = [max(1, n + pad) for n in self.n]
        self.dtype = dtype
        offa = np.cumsum([0] + [l * n for l, n in zip(self.lda, self.n)])
        offb = np.cumsum([0] + [l * r for l, r in zip(self.ldb, self.nrhs)])
        offp = np.cumsum([0] + [max(1, n) for n in self.n])
        self.offa, self.offb, self.offp = offa, offb, offp
        ha = np.full(int(offa[-1]) + 1, np.nan)
        hb = np.full(int(offb[-1]) + 1, np.nan)
        for s, (a, b) in enumerate(zip(mats, rhs)):
            n, l, lb = self.n[s], self.lda[s], self.ldb[s]
            if n:
                blk = np.full((n, l), np.nan)
                blk[:, :n] = a.T
                ha[offa[s]:offa[s + 1]] = blk.ravel()
                if self.nrhs[s]:
                    bb = np.full((self.nrhs[s], lb), np.nan)
                    bb[:, :n] = b.T
                    hb[offb[s]:offb[s + 1]] = bb.ravel()
        self.A = torch.from_numpy(ha).to(dtype).cuda()
        self.B = torch.from_numpy(hb).to(dtype).cuda()
        self.P = torch.zeros(int(offp[-1]) + 1, dtype=torch.int32, device="cuda")
        self.info = torch.full((len(mats),), -7, dtype=torch.int32, device="cuda")
        es = self.A.element_size()
        self.pa = torch.tensor([self.A.data_ptr() + int(o) * es for o in offa[:-1]], dtype=torch.int64).cuda()
        self.pb = torch.tensor([self.B.data_ptr() + int(o) * es for o in offb[:-1]], dtype=torch.int64).cuda()
        self.pp = torch.tensor([self.P.data_ptr() + int(o) * 4 for o in offp[:-1]], dtype=torch.int64).cuda()

    def run(self, dev, prec):
        dev.gesv_batched(prec, self.n, self.nrhs, self.pa, self.lda, self.pp, self.pb, self.ldb,
                         self.info)

    def x(self, s):
        n, r, lb = self.n[s], self.nrhs[s], self.ldb[s]
        v = self.B[int(self.offb[s]):int(self.offb[s + 1])].double().cpu().numpy()
        return v.reshape(r, lb)[:, :n].T.copy(order="F")

    def ipiv(self, s):
        return self.P[int(self.offp[s]):int(self.offp[s]) + self.n[s]].cpu().numpy()


def _systems(seed, sizes, shift_half=True):
    rng = np.random.default_rng(seed)
    mats, rhs = [], []
    for k, (n, r) in enumerate(sizes):
        a = rng.uniform(-1, 1, (n, n))
        if shift_half and k % 2 == 0:
            a += n * np.eye(n)
        mats.append(np.asfortranarray(a))
        rhs.append(np.asfortranarray(rng.standard_normal((n, r))))
    return mats, rhs


def _sresid(a, x, b, eps):
    if a.shape[0] == 0 or b.shape[1] == 0:
        return 0.0
    r = b - a @ x
    an = np.abs(a).sum(1).max()
    return max(np.abs(r[:, j]).max() / (an * max(np.abs(x[:, j]).max(), 1e-300) * a.shape[0] * eps)
               for j in range(b.shape[1]))


def _check_fp64(oracle, batch, mats, rhs):
    near_ties = 0
    for s, (a, b) in enumerate(zip(mats, rhs)):
        n = a.shape[0]
        if n == 0:
            continue
        lu = a.copy(order="F")
        piv, info, gap = oracle.pdgetrf(lu, max(n, 1), with_gap=True)
        assert info == 0
        got = batch.ipiv(s)
        if not np.array_equal(got, piv):
            first = int(np.argmax(got != piv))
            assert gap[first] < 1e-10, (s, n, first, gap[first])
            near_ties += 1
            continue
        if b.shape[1]:
            xo = oracle.pdgetrs(lu, piv, b.copy(order="F"), max(n, 1))
            x = batch.x(s)
            assert np.allclose(x, xo, rtol=0, atol=1e-10 * max(1.0, np.abs(xo).max())), (s, n)
            assert _sresid(a, x, b, 2.0 ** -53) < 16.0, (s, n)
    return near_ties


@pytest.mark.parametrize("w", [8, 16, 32])
def test_batched_fp64_matches_oracle(dev, oracle, w):
    rng = np.random.default_rng(11)
    ns = [1, 2, 3, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129, 255, 256, 257, 300, 511, 512] + \
         list(np.exp(rng.uniform(np.log(32), np.log(512), 40)).astype(int))
    sizes = [(int(n), int(rng.integers(1, 9))) for n in ns]
    _note(f"batched fp64 w={w} systems={len(sizes)}")
    mats, rhs = _systems(12, sizes)
    b = Batch(mats, rhs, pad=3)
    dev.set("batched_w", w)
    try:
        b.run(dev, "fp64")
    finally:
        dev.set("batched_w", 16)
    assert (b.info.cpu().numpy() == 0).all()
    ties = _check_fp64(oracle, b, mats, rhs)
    assert ties <= 2


def test_batched_edge_sizes_and_zero_pivot(dev, oracle):
    _note("batched edge sizes / zero pivot")
    sizes = [(0, 1), (5, 0), (1, 1), (4, 16), (700, 3), (1024, 1)]
    mats, rhs = _systems(21, sizes)
    # system 6: column 3 exactly zero -> INFO = 4 (1-based), B untouched
    z = np.asfortranarray(np.random.default_rng(22).uniform(-1, 1, (40, 40)))
    z[:, 3] = 0.0
    mats.append(z)
    rhs.append(np.asfortranarray(np.ones((40, 2))))
    b = Batch(mats, rhs)
    b.run(dev, "fp64")
    info = b.info.cpu().numpy()
    assert info.tolist() == [0, 0, 0, 0, 0, 0, 4]
    assert np.array_equal(b.x(6), rhs[6])
    lu = z.copy(order="F")
    piv, oinfo = oracle.pdgetrf(lu, 40)
    assert oinfo == 4
    _check_fp64(oracle, b, mats[:6], rhs[:6])


def test_batched_fp32(dev, oracle):
    rng = np.random.default_rng(31)
    sizes = [(int(n), int(rng.integers(1, 9))) for n in
             np.exp(rng.uniform(np.log(32), np.log(512), 60)).astype(int)]
    _note(f"batched fp32 systems={len(sizes)}")
    mats, rhs = _systems(32, sizes)
    mats32 = [a.astype(np.float32).astype(np.float64) for a in mats]
    b = Batch(mats32, rhs, dtype=torch.float32)
    b.run(dev, "fp32")
    assert (b.info.cpu().numpy() == 0).all()
    rhs32 = [r.astype(np.float32).astype(np.float64) for r in rhs]
    for s, (a, r) in enumerate(zip(mats32, rhs32)):
        assert _sresid(a, b.x(s), r, 2.0 ** -24) < 16.0, s


def test_batched_graph_replays_equal_eager(dev, oracle):
    rng = np.random.default_rng(41)
    sizes = [(int(n), int(rng.integers(1, 9))) for n in
             np.exp(rng.uniform(np.log(32), np.log(512), 80)).astype(int)]
    _note(f"batched graph replays systems={len(sizes)}")
    mats, rhs = _systems(42, sizes)
    b = Batch(mats, rhs)
    A0, B0 = b.A.clone(), b.B.clone()
    b.run(dev, "fp64")
    A_e, B_e, P_e = b.A.clone(), b.B.clone(), b.P.clone()
    dev.set("batched_graph", 1)
    try:
        for rep in range(3):
            b.A.copy_(A0)
            b.B.copy_(B0)
            b.P.zero_()
            torch.cuda.synchronize()
            b.run(dev, "fp64")
            # bit patterns (the padding rows hold NaN)
            bits = lambda t: t.view(torch.int64)  # noqa: E731
            assert torch.equal(bits(b.A), bits(A_e)) and torch.equal(bits(b.B), bits(B_e)), rep
            assert torch.equal(b.P, P_e), rep
            # eager work of the program between replays
            _ = (b.A * 2.0).sum().item()
    finally:
        dev.set("batched_graph", 0)
    _check_fp64(oracle, b, mats, rhs)


def test_batched_tuner(dev, oracle):
    rng = np.random.default_rng(51)
    ns = np.exp(rng.uniform(np.log(32), np.log(512), 300)).astype(int)
    nr = rng.integers(1, 9, 300)
    _note("batched tuner (300 systems)")
    w, k, ms = dev.gesv_batched_tune("fp64", ns.tolist(), nr.tolist())
    assert w in (8, 16, 32) and 1 <= k <= 4 and ms > 0
    assert dev.get("batched_w") == w and dev.get("batched_wgs_per_cu") == k
    _note(f"  tuned w={w} wgs_per_cu={k} ms={ms:.3f}")
    sizes = [(int(n), int(r)) for n, r in zip(ns[:40], nr[:40])]
    mats, rhs = _systems(52, sizes)
    b = Batch(mats, rhs)
    b.run(dev, "fp64")
    _check_fp64(oracle, b, mats, rhs)
    dev.set("batched_w", 16)
    dev.set("batched_wgs_per_cu", 2)
