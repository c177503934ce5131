"""The allocation-free batched entry (sla_gesv_batched_noalloc) against the
existing entry (sla_gesv_batched), through the C-ABI.

Both run the same persistent kernel on the same inputs, and every system's
factorization and solve is independent of the order in which the queue hands
systems out, so A (L\\U), IPIV, B (X) and INFO must agree bit for bit,
including the NaN padding between n and the leading dimensions.  Two calls in
a row (the second with fewer systems) check that the shared scratch array
carries nothing over; a third checks that a system with sizes out of range
gets INFO = -1 and is left untouched while the others still match.
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


def _systems(count, seed, nmax=300):
    rng = np.random.default_rng(seed)
    ns = rng.integers(0, nmax + 1, count)
    ns[: min(3, count)] = [nmax, 1, 0][: min(3, count)]
    mats, rhs = [], []
    for s, n in enumerate(ns):
        a = rng.uniform(-1, 1, (n, n))
        if s % 2 == 0:
            a += n * np.eye(n)
        mats.append(a)
        rhs.append(rng.uniform(-1, 1, (n, int(rng.integers(0, 17)))))
    return mats, rhs


class Packed:
    """One device buffer each for A, B, IPIV (NaN padding, lda = n + 3) and
    the pointer / size arrays of both entries."""

    def __init__(self, mats, rhs, dtype):
        self.n = [a.shape[0] for a in mats]
        self.nrhs = [b.shape[1] for b in rhs]
        self.lda = [n + 3 for n in self.n]
        self.ldb = [n + 3 for n in self.n]
        offa = np.cumsum([0] + [l * n for l, n in zip(self.lda, self.n)])
        offb = np.cumsum([0] + [l * r for l, r in zip(self.ldb, self.nrhs)])
        offp = np.cumsum([0] + [max(1, n) for n in self.n])
        ha = np.full(int(offa[-1]) + 1, np.nan)
        hb = np.full(int(offb[-1]) + 1, np.nan)
        for s, (a, b) in enumerate(zip(mats, rhs)):
            n, la, lb = self.n[s], self.lda[s], self.ldb[s]
            if n:
                blk = np.full((n, la), np.nan)
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
        self.pa = torch.tensor([self.A.data_ptr() + int(o) * es for o in offa[:-1]],
                               dtype=torch.int64).cuda()
        self.pb = torch.tensor([self.B.data_ptr() + int(o) * es for o in offb[:-1]],
                               dtype=torch.int64).cuda()
        self.pp = torch.tensor([self.P.data_ptr() + int(o) * 4 for o in offp[:-1]],
                               dtype=torch.int64).cuda()
        dint = lambda v: torch.tensor(v, dtype=torch.int32).cuda()  # noqa: E731
        self.dn, self.dnrhs = dint(self.n), dint(self.nrhs)
        self.dlda, self.dldb = dint(self.lda), dint(self.ldb)

    def bits(self):
        """Raw storage of every output, NaNs compared by their bits."""
        ib = torch.int64 if self.A.dtype == torch.float64 else torch.int32
        return (self.A.view(ib).cpu(), self.B.view(ib).cpu(), self.P.cpu(), self.info.cpu())


def _both(dev, prec, mats, rhs):
    dt = torch.float64 if prec == "fp64" else torch.float32
    ref, new = Packed(mats, rhs, dt), Packed(mats, rhs, dt)
    dev.gesv_batched(prec, ref.n, ref.nrhs, ref.pa, ref.lda, ref.pp, ref.pb, ref.ldb, ref.info)
    dev.gesv_batched_noalloc(prec, len(new.n), new.dn, new.dnrhs, new.pa, new.dlda, new.pp,
                             new.pb, new.dldb, new.info)
    torch.cuda.synchronize()
    return ref, new


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_batched_twice(dev, prec):
    for call, (count, seed) in enumerate([(400, 1), (37, 2)]):
        _note(f"noalloc {prec} call {call}: {count} systems")
        mats, rhs = _systems(count, seed + (0 if prec == "fp64" else 10))
        ref, new = _both(dev, prec, mats, rhs)
        for name, r, g in zip(("A", "B", "IPIV", "INFO"), ref.bits(), new.bits()):
            assert torch.equal(r, g), f"{name} differs from sla_gesv_batched (call {call})"
        assert int((ref.info != 0).sum()) == 0


def test_noalloc_rejects_bad_sizes_per_system(dev):
    _note("noalloc: a system with n > 1024 and one with lda < n")
    mats, rhs = _systems(20, 3, nmax=64)
    good = Packed(mats, rhs, torch.float64)
    bad = Packed(mats, rhs, torch.float64)
    nv, lv = list(bad.n), list(bad.lda)
    nv[5], lv[9] = 2000, max(1, bad.n[9] - 1)
    bad.dn = torch.tensor(nv, dtype=torch.int32).cuda()
    bad.dlda = torch.tensor(lv, dtype=torch.int32).cuda()
    a5 = bad.A.clone()
    dev.gesv_batched(("fp64"), good.n, good.nrhs, good.pa, good.lda, good.pp, good.pb, good.ldb,
                     good.info)
    dev.gesv_batched_noalloc("fp64", 20, bad.dn, bad.dnrhs, bad.pa, bad.dlda, bad.pp, bad.pb,
                             bad.dldb, bad.info)
    torch.cuda.synchronize()
    info = bad.info.cpu().numpy()
    assert info[5] == -1 and info[9] == -1
    ok = [s for s in range(20) if s not in (5, 9)]
    assert np.array_equal(info[ok], good.info.cpu().numpy()[ok])
    # the rejected systems' storage is untouched; the others match bit for bit
    gA, bA = good.A.view(torch.int64).cpu(), bad.A.view(torch.int64).cpu()
    a5b = a5.view(torch.int64).cpu()
    offa = np.cumsum([0] + [l * n for l, n in zip(good.lda, good.n)])
    for s in range(20):
        sl = slice(int(offa[s]), int(offa[s + 1]))
        if s in (5, 9):
            assert torch.equal(bA[sl], a5b[sl])
        else:
            assert torch.equal(bA[sl], gA[sl])


def test_noalloc_batch_limit(dev):
    from scalapack_amd.device import SlaError
    z = torch.zeros(1, dtype=torch.int32, device="cuda")
    p = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(SlaError, match="batch outside"):
        dev.gesv_batched_noalloc("fp64", 65537, z, z, p, z, p, p, z, z)
