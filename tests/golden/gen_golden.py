"""Writes the golden fixtures of the LU path from the reference itself.

Runs oracle/_ref/ref_lu (the reference's PDGETRF/PDGETRS compiled from its
sources by oracle/Makefile, host BLAS = MKL, sequential) and stores, per case,
IPIV, INFO, the solution X and the tester's scaled residual.  Cases:

  * LU.dat (TESTING/traditional/LU.dat:5-16): (M, N) in (4,4) (10,12) (17,13)
    (13,13), NB in 2 3 4, NRHS in 1 3 9, grids 1x1, 2x2, 1x4, 4x1, matrices
    from PDMATGEN with IASEED = 100 and B with IBSEED = 200
    (pdludriver.f:130-131);
  * config 1: N = 1024, NB = 64, NRHS = 1, b = A * ones, grids 1x1 and 2x2;
  * config 2: N = 16384, NB = 256, NRHS = 16, B i.i.d. standard normal from
    numpy.random.default_rng(200) (written column-major), grid 1x1;
  * the 6 x 6 example of EXAMPLE/DSCAEXMAT.dat / DSCAEXRHS.dat with its exact
    rational solution;
  * the routines beside the solve (ref_lu aux): PDLANGE 'M' '1' 'I' 'F' of A
    and of A(2:N-2, 3:N), PDGECON '1' and 'I', PDGETRS 'T', and PDGERFS 'N'
    from the PDGETRS solution, 'N' and 'T' from perturbed starts, for
    (N, NB, NRHS) in AUX_CASES with A = PDMATGEN(100), B = PDMATGEN(200);
  * equilibration and the expert driver (ref_lu svx): PDGEEQU, PDLAQGE and
    PDGESVX (FACT E/TRANS N, E/T, N/N) on PDMATGEN(100) with exact
    power-of-two row and column scales, for SVX_CASES;
  * ASCII matrix I/O (ref_lu io): the reference's PDLAWRITE file of a 9 x 6
    matrix with special values, byte for byte, and its PDLAREAD of
    EXAMPLE/DSCAEXMAT.dat.

Usage (in the build container, where /root/reference exists):
    make -C oracle ref && python tests/golden/gen_golden.py [--skip-c2]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF_LU = ROOT / "oracle" / "_ref" / "ref_lu"
MPIEXEC = "/opt/conda/bin/mpiexec"
REFERENCE = Path("/root/reference")

LU_DAT_MN = [(4, 4), (10, 12), (17, 13), (13, 13)]
LU_DAT_NB = [2, 3, 4]
LU_DAT_NRHS = [1, 3, 9]
LU_DAT_GRIDS = [(1, 1), (2, 2), (1, 4), (4, 1)]
AUX_CASES = [(37, 4, 2), (200, 32, 3), (513, 64, 1)]
SVX_CASES = [(48, 8, 2), (150, 32, 3)]


def read_out(path: Path) -> dict:
    b = path.read_bytes()
    hdr = np.frombuffer(b[:40], np.int32)
    M, N, NB, NRHS, P, Q, i1, i2, fl = (int(x) for x in hdr[1:10])
    dh = np.frombuffer(b[40:72], np.float64)
    off = 72
    mn = min(M, N)
    ipiv = np.frombuffer(b[off:off + 4 * mn], np.int32).copy()
    off += 4 * mn
    x = None
    if fl & 1:
        x = np.frombuffer(b[off:off + 8 * N * NRHS], np.float64).reshape(NRHS, N).T.copy(order="F")
    return dict(M=M, N=N, NB=NB, NRHS=NRHS, P=P, Q=Q, info_getrf=i1, info_getrs=i2,
                t_getrf=dh[0], t_getrs=dh[1], sresid=dh[2], anorm=dh[3], ipiv=ipiv, x=x)


def run_ref(M, N, NB, NRHS, P, Q, bmode, threads: int = 1) -> dict:
    env = dict(os.environ)
    env["MKL_THREADING_LAYER"] = "SEQUENTIAL" if threads == 1 else "GNU"
    env["MKL_NUM_THREADS"] = str(threads)
    env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "o.bin"
        cmd = [str(REF_LU), str(M), str(N), str(NB), str(NRHS), str(P), str(Q), "100", bmode,
               str(out)]
        if P * Q > 1:
            cmd = [MPIEXEC, "-n", str(P * Q)] + cmd
        subprocess.run(cmd, check=True, env=env, capture_output=True)
        return read_out(out)


def run_aux(N, NB, NRHS) -> dict:
    """ref_lu aux: see oracle/ref_driver.c aux() for the file layout."""
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "aux.bin"
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
        subprocess.run([str(REF_LU), "aux", str(N), str(NB), str(NRHS), "100", "200", str(out)],
                       check=True, capture_output=True, env=env)
        b = out.read_bytes()
    hdr = np.frombuffer(b[:44], np.int32)
    assert hdr[0] == 0x41555852 and tuple(hdr[1:4]) == (N, NB, NRHS)
    off = 44
    dh = np.frombuffer(b[off:off + 80], np.float64)
    off += 80
    ipiv = np.frombuffer(b[off:off + 4 * N], np.int32).copy()
    off += 4 * N
    nr = N * NRHS

    def mat():
        nonlocal off
        m = np.frombuffer(b[off:off + 8 * nr], np.float64).reshape(NRHS, N).T.copy(order="F")
        off += 8 * nr
        return m

    def vec():
        nonlocal off
        v = np.frombuffer(b[off:off + 8 * NRHS], np.float64).copy()
        off += 8 * NRHS
        return v

    r = dict(info=hdr[4:11].copy(), norms=dh[0:4].copy(), subnorms=dh[4:8].copy(),
             rcond=dh[8:10].copy(), ipiv=ipiv, xT=mat())
    for c in range(3):
        r[f"rfs{c}_x0"] = mat()
        r[f"rfs{c}_x"] = mat()
        r[f"rfs{c}_ferr"] = vec()
        r[f"rfs{c}_berr"] = vec()
    assert off == len(b)
    return r


def run_svx(N, NB, NRHS) -> dict:
    """ref_lu svx: see oracle/ref_driver.c svx() for the file layout."""
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "svx.bin"
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL", MKL_NUM_THREADS="1")
        subprocess.run([str(REF_LU), "svx", str(N), str(NB), str(NRHS), str(out)], check=True,
                       capture_output=True, env=env)
        b = out.read_bytes()
    hdr = np.frombuffer(b[:24], np.int32)
    assert hdr[0] == 0x53565852 and tuple(hdr[1:4]) == (N, NB, NRHS)
    cinfo = np.frombuffer(b[24:48], np.int32)
    off = 48
    d = np.frombuffer(b[off:], np.float64)
    nr = N * NRHS
    q = 0

    def take(k):
        nonlocal q
        v = d[q:q + k].copy()
        q += k
        return v

    r = dict(info_equ=np.array(hdr[4]), equed_laqge=np.array(hdr[5]))
    r["rowcnd"], r["colcnd"], r["amax"] = take(1), take(1), take(1)
    r["R"], r["C"] = take(N), take(N)
    r["A_laqge"] = take(N * N).reshape(N, N).T.copy(order="F")
    for c in range(3):
        r[f"c{c}_info"] = np.array(cinfo[2 * c])
        r[f"c{c}_equed"] = np.array(cinfo[2 * c + 1])
        r[f"c{c}_rcond"] = take(1)
        r[f"c{c}_R"], r[f"c{c}_C"] = take(N), take(N)
        r[f"c{c}_X"] = take(nr).reshape(NRHS, N).T.copy(order="F")
        r[f"c{c}_ferr"], r[f"c{c}_berr"] = take(NRHS), take(NRHS)
        r[f"c{c}_B"] = take(nr).reshape(NRHS, N).T.copy(order="F")
    assert q == d.size
    return r


def run_io() -> None:
    """ref_lu io: the reference's PDLAWRITE of a 9 x 6 matrix with special
    values (kept byte for byte as pdlawrite_9x6.txt) and its PDLAREAD of
    EXAMPLE/DSCAEXMAT.dat into a 12 x 8 array (pdlaread_dscaexmat.npy)."""
    with tempfile.TemporaryDirectory() as td:
        txt, raw = Path(td) / "w.txt", Path(td) / "r.bin"
        env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL")
        subprocess.run([str(REF_LU), "io", "9", "6", "4", str(txt),
                        str(REFERENCE / "EXAMPLE" / "DSCAEXMAT.dat"), str(raw)], check=True,
                       capture_output=True, env=env, cwd=td)
        (HERE / "pdlawrite_9x6.txt").write_bytes(txt.read_bytes())
        r = np.fromfile(raw, np.float64).reshape(8, 12).T.copy(order="F")
        np.save(HERE / "pdlaread_dscaexmat.npy", r, allow_pickle=False)


def read_dat(path: Path) -> np.ndarray:
    """PDLAREAD's format (TOOLS/pdlaread.f:67-116): 'M N', then values down
    the columns, one per line, Fortran D exponents."""
    lines = path.read_text().split("\n")
    m, n = (int(v) for v in lines[0].split()[:2])
    vals = [float(t.replace("D", "E")) for t in " ".join(lines[1:]).split()]
    return np.array(vals[: m * n], dtype=np.float64).reshape(n, m).T.copy(order="F")


def exact_solve(a: np.ndarray, b: np.ndarray) -> list[Fraction]:
    n = a.shape[0]
    m = [[Fraction(float(a[i, j])) for j in range(n)] + [Fraction(float(b[i]))] for i in range(n)]
    for c in range(n):
        p = max(range(c, n), key=lambda r: abs(m[r][c]))
        m[c], m[p] = m[p], m[c]
        for r in range(n):
            if r != c and m[r][c] != 0:
                f = m[r][c] / m[c][c]
                m[r] = [x - f * y for x, y in zip(m[r], m[c])]
    return [m[i][n] / m[i][i] for i in range(n)]


def _desc(m=4, n=4, mb=2, nb=2, rsrc=0, csrc=0, lld=4, ctxt=0, dtype=1):
    return [dtype, ctxt, m, n, mb, nb, rsrc, csrc, lld]


def argcheck_cases() -> list[dict]:
    """Argument lists exercising every check of PDGETRF / PDGETRS / PDGESV
    (chk1mat.f and the routines' own checks)."""
    c = []
    f = lambda **k: c.append(dict(routine="pdgetrf", **k))  # noqa: E731
    f(args=[4, 4, 1, 1], desca=_desc())
    f(args=[4, 4, 1, 1], desca=_desc(dtype=2))
    f(args=[4, 4, 1, 1], desca=_desc(ctxt=-999))
    f(args=[-1, 4, 1, 1], desca=_desc())
    f(args=[4, -1, 1, 1], desca=_desc())
    f(args=[4, 4, 0, 1], desca=_desc())
    f(args=[4, 4, 1, 0], desca=_desc())
    f(args=[4, 4, 1, 1], desca=_desc(mb=0))
    f(args=[4, 4, 1, 1], desca=_desc(nb=0))
    f(args=[4, 4, 1, 1], desca=_desc(rsrc=1))
    f(args=[4, 4, 1, 1], desca=_desc(csrc=-1))
    f(args=[4, 4, 1, 1], desca=_desc(lld=0))
    f(args=[4, 4, 1, 1], desca=_desc(lld=3))
    f(args=[4, 4, 1, 1], desca=_desc(m=0))
    f(args=[4, 4, 5, 1], desca=_desc())
    f(args=[4, 4, 3, 1], desca=_desc())
    f(args=[4, 4, 1, 3], desca=_desc())
    f(args=[2, 2, 2, 1], desca=_desc())
    f(args=[2, 2, 1, 2], desca=_desc())
    f(args=[4, 4, 1, 1], desca=_desc(mb=2, nb=4))
    f(args=[0, 4, 1, 1], desca=_desc())
    f(args=[1, 3, 1, 1], desca=_desc(m=1, n=3, lld=1))
    g = lambda **k: c.append(dict(routine="pdgetrs", **k))  # noqa: E731
    g(trans="N", args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    g(trans="X", args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[2, 1, 2, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[2, 1, 1, 2], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[4, 1, 1, 1], desca=_desc(mb=4), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[2, 1, 1, 1], desca=_desc(), bargs=[2, 1], descb=_desc(n=1))
    g(trans="N", args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1, mb=4, nb=4))
    g(trans="N", args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1, lld=2))
    g(trans="N", args=[4, 2, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[4, 1, 1, 1], desca=_desc(ctxt=-999), bargs=[1, 1], descb=_desc(n=1))
    g(trans="N", args=[0, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    h = lambda **k: c.append(dict(routine="pdgesv", **k))  # noqa: E731
    h(args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    h(args=[4, 1, 1, 1], desca=_desc(dtype=7), bargs=[1, 1], descb=_desc(n=1))
    h(args=[2, 1, 2, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    h(args=[2, 1, 1, 2], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    h(args=[4, 1, 1, 1], desca=_desc(nb=4), bargs=[1, 1], descb=_desc(n=1))
    h(args=[2, 1, 1, 1], desca=_desc(), bargs=[3, 1], descb=_desc(n=1))
    h(args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1, mb=1, nb=1))
    h(args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1, lld=3))
    h(args=[4, -2, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1))
    h(args=[4, 1, 1, 1], desca=_desc(), bargs=[1, 1], descb=_desc(n=1, ctxt=-999))
    return c


def run_argcheck(case: dict) -> int:
    r = case["routine"]
    if r == "pdgetrf":
        argv = [*case["args"], *case["desca"]]
    else:
        n, nrhs, ia, ja = case["args"]
        argv = [n, nrhs, ia, ja, *case["desca"], *case["bargs"], *case["descb"]]
        if r == "pdgetrs":
            argv = [case["trans"], *argv]
    out = subprocess.run([str(REF_LU), "argcheck", r, *map(str, argv)], check=True,
                         capture_output=True, text=True).stdout
    return int(out.split("INFO")[1].split()[0])


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--skip-c2", action="store_true")
    ap.add_argument("--only-aux", action="store_true",
                    help="write only aux_cases.npz (the routines beside the solve)")
    args = ap.parse_args()
    if not REF_LU.exists():
        sys.exit(f"{REF_LU} missing: run `make -C oracle ref` first")
    if args.only_aux:
        aux = {}
        for (N, NB, NRHS) in AUX_CASES:
            for k, v in run_aux(N, NB, NRHS).items():
                aux[f"N{N}_NB{NB}_NRHS{NRHS}_{k}"] = v
        np.savez_compressed(HERE / "aux_cases.npz", **aux)
        print("aux:", AUX_CASES)
        svx = {}
        for (N, NB, NRHS) in SVX_CASES:
            for k, v in run_svx(N, NB, NRHS).items():
                svx[f"N{N}_NB{NB}_NRHS{NRHS}_{k}"] = v
        np.savez_compressed(HERE / "svx_cases.npz", **svx)
        print("svx:", SVX_CASES)
        run_io()
        print("io: pdlawrite_9x6.txt, pdlaread_dscaexmat.npy")
        return

    # LU.dat cases
    lu = {}
    for (M, N) in LU_DAT_MN:
        for NB in LU_DAT_NB:
            for (P, Q) in LU_DAT_GRIDS:
                for NRHS in (LU_DAT_NRHS if M == N else [1]):
                    r = run_ref(M, N, NB, NRHS, P, Q, "lcg:200")
                    key = f"M{M}_N{N}_NB{NB}_NRHS{NRHS}_P{P}_Q{Q}"
                    lu[key + "_ipiv"] = r["ipiv"]
                    lu[key + "_info"] = np.array([r["info_getrf"], r["info_getrs"]], np.int32)
                    lu[key + "_sresid"] = np.array([r["sresid"]])
                    if r["x"] is not None:
                        lu[key + "_x"] = r["x"]
    np.savez_compressed(HERE / "lu_dat_cases.npz", **lu)
    print("lu_dat_cases:", len(lu) // 3, "cases")

    # argument checks
    cases = argcheck_cases()
    for case in cases:
        case["info"] = run_argcheck(case)
    (HERE / "argcheck_cases.json").write_text(json.dumps(cases, indent=1) + "\n")
    print("argcheck:", [c["info"] for c in cases])

    # config 1
    c1 = {}
    for (P, Q) in [(1, 1), (2, 2)]:
        r = run_ref(1024, 1024, 64, 1, P, Q, "ones")
        k = f"P{P}_Q{Q}"
        c1[k + "_ipiv"] = r["ipiv"]
        c1[k + "_x"] = r["x"]
        c1[k + "_info"] = np.array([r["info_getrf"], r["info_getrs"]], np.int32)
        c1[k + "_sresid"] = np.array([r["sresid"]])
    np.savez_compressed(HERE / "c1_n1024_nb64.npz", **c1)
    print("c1: ipiv 1x1 == 2x2:", np.array_equal(c1["P1_Q1_ipiv"], c1["P2_Q2_ipiv"]))

    # routines beside the solve
    aux = {}
    for (N, NB, NRHS) in AUX_CASES:
        r = run_aux(N, NB, NRHS)
        for k, v in r.items():
            aux[f"N{N}_NB{NB}_NRHS{NRHS}_{k}"] = v
    np.savez_compressed(HERE / "aux_cases.npz", **aux)
    print("aux:", AUX_CASES)
    svx = {}
    for (N, NB, NRHS) in SVX_CASES:
        for k, v in run_svx(N, NB, NRHS).items():
            svx[f"N{N}_NB{NB}_NRHS{NRHS}_{k}"] = v
    np.savez_compressed(HERE / "svx_cases.npz", **svx)
    print("svx:", SVX_CASES)
    run_io()

    # 6 x 6 example
    a6 = read_dat(REFERENCE / "EXAMPLE" / "DSCAEXMAT.dat")
    b6 = read_dat(REFERENCE / "EXAMPLE" / "DSCAEXRHS.dat")
    xs = exact_solve(a6, b6[:, 0])
    np.savez_compressed(HERE / "example_6x6.npz", a=a6, b=b6,
                        x_num=np.array([x.numerator for x in xs], np.int64),
                        x_den=np.array([x.denominator for x in xs], np.int64))
    print("6x6 exact x:", [str(x) for x in xs])

    # config 2
    if not args.skip_c2:
        n, nrhs = 16384, 16
        rng = np.random.default_rng(200)
        B = np.asfortranarray(rng.standard_normal((n, nrhs)))
        with tempfile.TemporaryDirectory() as td:
            bf = Path(td) / "b.bin"
            B.T.tofile(bf)  # column-major
            r = run_ref(n, n, 256, nrhs, 1, 1, f"file:{bf}", threads=os.cpu_count() or 1)
        np.savez_compressed(HERE / "c2_n16384_nb256.npz", ipiv=r["ipiv"],
                            info=np.array([r["info_getrf"], r["info_getrs"]], np.int32),
                            sresid=np.array([r["sresid"]]), x=r["x"],
                            t=np.array([r["t_getrf"], r["t_getrs"]]))
        print("c2: sresid", r["sresid"], "t_getrf", r["t_getrf"])


if __name__ == "__main__":
    main()
