"""The drop-in boundary without a GPU: the C-ABI library loads, exports every
symbol the public headers declare, its BLACS 1 x 1 / TOOLS helpers follow the
reference's closed forms, and every illegal-argument INFO code equals the one
the reference itself returns (tests/golden/argcheck_cases.json, written by
running the reference's PDGETRF / PDGETRS / PDGESV)."""
import ctypes

import numpy as np
import pytest

import scalapack_amd
from scalapack_amd import api
from scalapack_amd._lib import declared_symbols


def test_library_loads_and_exports_declared_symbols():
    lib = scalapack_amd.load()
    shim = scalapack_amd._lib.shim()
    decl = declared_symbols()
    assert len(decl["scalapack_amd.h"]) >= 15
    for name in decl["scalapack_amd.h"]:
        assert hasattr(lib, name), name
    for name in decl["sla_blacs1x1.h"]:
        assert hasattr(shim, name), name


def test_reference_fortran_entry_points_are_declared():
    decl = declared_symbols()["scalapack_amd.h"]
    for name in ("pdgesv_", "pdgetrf_", "pdgetrs_", "pdgetf2_", "pdlaswp_", "pdlange_",
                 "pdgecon_", "pdgerfs_"):
        assert name in decl


def test_blacs_1x1_context():
    assert api.blacs_pinfo() == (0, 1)
    ctx = api.blacs_gridinit(1, 1)
    assert api.blacs_gridinfo(ctx) == (1, 1, 0, 0)
    api.blacs_gridexit(ctx)
    assert api.blacs_gridinfo(ctx) == (-1, -1, -1, -1)
    bad = api.blacs_gridinit(2, 2)   # needs 4 processes
    assert api.blacs_gridinfo(bad)[0] == -1


def numroc_closed(n, nb, iproc, isrc, nprocs):
    # count the entries of 0..n-1 owned by iproc, by brute force
    return sum(1 for g in range(n) if (isrc + g // nb) % nprocs == iproc)


@pytest.mark.parametrize("n,nb,np_", [(10, 3, 1), (10, 3, 2), (17, 4, 3), (1024, 64, 2), (5, 8, 4)])
def test_numroc_matches_enumeration(n, nb, np_):
    for p in range(np_):
        for src in range(np_):
            assert api.numroc(n, nb, p, src, np_) == numroc_closed(n, nb, p, src, np_)


def test_descinit():
    ctx = api.blacs_gridinit(1, 1)
    d, info = api.descinit(1024, 1024, 64, 64, 0, 0, ctx, 1024)
    assert info == 0
    assert d.tolist() == [1, ctx, 1024, 1024, 64, 64, 0, 0, 1024]
    d, info = api.descinit(100, 5, 8, 8, 0, 0, ctx, 50)   # LLD < M
    assert info == -9


def _run_argcheck(case, ctx):
    def fix(d):
        d = np.array(d, np.int32)
        d[1] = ctx + 77 if d[1] == -999 else ctx
        return d
    # identity A laid out with the descriptor's LLD, as the reference driver
    # builds it (oracle/ref_driver.c, argcheck mode)
    lld = max(1, case["desca"][8])
    a = np.zeros(max(64, lld * 32))
    for i in range(32):
        if i * (lld + 1) < a.size:
            a[i * (lld + 1)] = 1.0
    a = np.asfortranarray(a.reshape(-1, lld).T) if lld > 1 else np.asfortranarray(a[None, :])
    b = np.asfortranarray(np.zeros((8, 2)))
    ipiv = np.arange(1, 33, dtype=np.int32)
    r = case["routine"]
    if r == "pdgetrf":
        m, n, ia, ja = case["args"]
        return api.pdgetrf(m, n, a, ia, ja, fix(case["desca"]), ipiv)
    n, nrhs, ia, ja = case["args"]
    ib, jb = case["bargs"]
    if r == "pdgetrs":
        return api.pdgetrs(case["trans"], n, nrhs, a, ia, ja, fix(case["desca"]), ipiv, b, ib,
                           jb, fix(case["descb"]))
    return api.pdgesv(n, nrhs, a, ia, ja, fix(case["desca"]), ipiv, b, ib, jb,
                      fix(case["descb"]))


def _cpu_safe(case):
    """Cases whose INFO is decided before any device work."""
    if case["info"] < 0:
        return True
    args = case["args"]
    if case["routine"] == "pdgetrf":
        return args[0] == 0 or args[1] == 0 or case["desca"][2] == 1
    return args[0] == 0


def test_illegal_argument_info_matches_reference(golden, capfd):
    ctx = api.blacs_gridinit(1, 1)
    checked = 0
    for case in golden.argcheck_cases():
        if not _cpu_safe(case):
            continue
        got = _run_argcheck(case, ctx)
        assert got == case["info"], case
        checked += 1
    assert checked >= 35
    out = capfd.readouterr().out
    assert "On entry to PDGETRF parameter number  606 had an illegal value" in out


def test_pdgetrs_illegal_trans_is_reported(capfd):
    """TRANS other than N / T / C is argument 1 (pdgetrs.f:208-210); 'T'
    and 'C' pass the checks and run on the GPU (tests/test_gpu_aux.py)."""
    ctx = api.blacs_gridinit(1, 1)
    d, _ = api.descinit(4, 4, 2, 2, 0, 0, ctx, 4)
    db, _ = api.descinit(4, 1, 2, 2, 0, 0, ctx, 4)
    a = np.asfortranarray(np.eye(4))
    b = np.asfortranarray(np.ones((4, 1)))
    assert api.pdgetrs("X", 4, 1, a, 1, 1, d, np.arange(1, 5, dtype=np.int32), b, 1, 1, db) == -1
    assert "On entry to PDGETRS parameter number    1 had an illegal value" in capfd.readouterr().out


def test_aux_entry_argument_checks(capfd):
    """PDGECON / PDGERFS argument checks and workspace queries return before
    any GPU work (pdgecon.f:215-285, pdgerfs.f:308-404)."""
    ctx = api.blacs_gridinit(1, 1)
    n, nb = 8, 4
    d, _ = api.descinit(n, n, nb, nb, 0, 0, ctx, n)
    db, _ = api.descinit(n, 2, nb, nb, 0, 0, ctx, n)
    a = np.asfortranarray(np.eye(n))
    work, iwork = np.zeros(64), np.zeros(16, np.int32)
    rc, info = api.pdgecon("1", n, a, 1, 1, d, 1.0, work, -1, iwork, 1)
    assert info == 0 and work[0] == 5 * n + nb and iwork[0] == n   # LWMIN, LIWMIN on 1 x 1
    assert api.pdgecon("Q", n, a, 1, 1, d, 1.0, work, 64, iwork, 16)[1] == -1
    assert api.pdgecon("I", n, a, 1, 1, d, -2.0, work, 64, iwork, 16)[1] == -7
    assert api.pdgecon("I", n, a, 1, 1, d, 1.0, work, 5 * n + nb - 1, iwork, 16)[1] == -10
    assert api.pdgecon("I", n, a, 1, 1, d, 1.0, work, 64, iwork, n - 1)[1] == -12
    assert api.pdgecon("I", n, a, 2, 1, d, 1.0, work, 64, iwork, 16)[1] == -2  # N+IA-1 > M_A (CHK1MAT)
    ipiv = np.arange(1, n + 1, dtype=np.int32)
    b = np.asfortranarray(np.ones((n, 2)))
    x = b.copy(order="F")
    fe, be = np.zeros(2), np.zeros(2)

    def rfs(trans="N", lwork=3 * n, liwork=n, ib=1, descx=db, ia=1):
        return api.pdgerfs(trans, n if ia == 1 else n - nb, 2, a, ia, 1, d, a, ia, 1, d, ipiv, b,
                           ib, 1, db, x, 1, 1, descx, fe, be, work, lwork, iwork, liwork)
    assert rfs(lwork=-1) == 0 and work[0] == 3 * n and iwork[0] == n
    assert rfs("Z") == -1
    assert rfs(lwork=3 * n - 1) == -24
    assert rfs(liwork=n - 1) == -26
    bad_x = db.copy()
    bad_x[4] = nb + 1
    assert rfs(descx=bad_x) == -(2000 + 5)     # MB_B != MB_X
    assert rfs(ia=3) == -5                     # IA not aligned to MB_A
    out = capfd.readouterr().out
    assert "On entry to PDGECON parameter number    1 had an illegal value" in out
    assert "On entry to PDGERFS parameter number   24 had an illegal value" in out


def test_python_mirror_rejects_c_ordered_arrays():
    ctx = api.blacs_gridinit(1, 1)
    d, _ = api.descinit(4, 4, 2, 2, 0, 0, ctx, 4)
    with pytest.raises(TypeError):
        api.pdgetrf(4, 4, np.eye(4)[:, ::-1].copy(order="C"), 1, 1, d, np.zeros(6, np.int32))


def test_device_handle_params_need_no_gpu_for_validation():
    lib = scalapack_amd.load()
    # an unknown parameter name is rejected by the C-ABI before any device work
    rc = lib.sla_set_param(ctypes.c_void_p(0), b"no_such_knob", 1)
    assert rc != 0
    assert "unknown parameter" in scalapack_amd._lib.last_error()


def test_pdgesvx_argument_checks(capfd):
    """PDGESVX's checks and workspace query return before any GPU work
    (pdgesvx.f:452-620): LWMIN = max(5N + NB, 5N + ceil(N/NB) NB) on 1 x 1."""
    ctx = api.blacs_gridinit(1, 1)
    n, nb, nrhs = 10, 4, 1
    d, _ = api.descinit(n, n, nb, nb, 0, 0, ctx, n)
    db, _ = api.descinit(n, nrhs, nb, nb, 0, 0, ctx, n)
    a = np.asfortranarray(np.eye(n))
    af = np.asfortranarray(np.zeros((n, n)))
    b = np.asfortranarray(np.ones((n, nrhs)))
    x = np.asfortranarray(np.zeros((n, nrhs)))
    ipiv = np.zeros(n + nb, np.int32)
    r, c = np.ones(n), np.ones(n)
    fe, be = np.zeros(nrhs), np.zeros(nrhs)
    work, iwork = np.zeros(200), np.zeros(20, np.int32)

    def svx(fact="N", trans="N", lwork=200, liwork=20, equed="N", rr=r, ja=1):
        return api.pdgesvx(fact, trans, n, nrhs, a, 1, ja, d, af, 1, 1, d, ipiv, equed, rr, c, b,
                           1, 1, db, x, 1, 1, db, fe, be, work, lwork, iwork, liwork)[0]
    assert svx(lwork=-1) == 0
    lwmin = max(5 * n + nb, 5 * n + -(-n // nb) * nb)
    assert work[0] == lwmin and iwork[0] == n
    assert svx(fact="Q") == -1
    assert svx(trans="Q") == -2
    assert svx(fact="F", equed="Z") == -13
    assert svx(fact="F", equed="R", rr=np.zeros(n)) == -14
    assert svx(lwork=lwmin - 1) == -29
    assert svx(liwork=n - 1) == -31
    out = capfd.readouterr().out
    assert "On entry to PDGESVX parameter number   29 had an illegal value" in out


_PDGETF2_ABORT = """
import sys
sys.path.insert(0, {root!r})
import numpy as np
from scalapack_amd import api
ctx = api.blacs_gridinit(1, 1)
m, n, nb, ja = {m}, {n}, {nb}, {ja}
d, _ = api.descinit(m, ja - 1 + n + 1, nb, nb, 0, 0, ctx, m)
a = np.asfortranarray(np.eye(m, ja - 1 + n + 1))
ipiv = np.zeros(m + nb, np.int32)
api.pdgetf2(m, n, a, 1, ja, d, ipiv)
print("RETURNED")
"""


@pytest.mark.parametrize("n,nb,ja,pos", [(40, 32, 1, 2), (32, 32, 2, 2), (16, 32, 2, 5)])
def test_pdgetf2_illegal_argument_aborts(n, nb, ja, pos):
    """PDGETF2 reports an illegal argument through PXERBLA and then calls
    BLACS_ABORT (SRC/pdgetf2.f:177-193): N + ICOFF > NB is argument 2, a JA
    off a block boundary argument 5.  The process ends with the abort's
    error code; the call never returns.  Runs in a child process (no GPU
    work happens before the abort)."""
    import subprocess
    import sys
    from pathlib import Path
    root = str(Path(__file__).resolve().parents[1])
    code = _PDGETF2_ABORT.format(root=root, m=64, n=n, nb=nb, ja=ja)
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 1, (proc.returncode, proc.stdout, proc.stderr)
    assert "RETURNED" not in proc.stdout
    assert f"On entry to PDGETF2 parameter number {pos:4d} had an illegal value" in proc.stdout
    assert "exiting with error #1" in proc.stderr
