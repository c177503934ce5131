"""Readers of the golden fixtures written by tests/golden/gen_golden.py from
the reference's own PDGETRF/PDGETRS (see that script for the cases)."""
from __future__ import annotations

import json
from fractions import Fraction
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent / "golden"


def lu_dat() -> dict:
    z = np.load(HERE / "lu_dat_cases.npz")
    return {k: z[k] for k in z.files}


def lu_dat_cases():
    """(M, N, NB, NRHS, P, Q) of every stored LU.dat case."""
    out = set()
    for k in lu_dat():
        p = k.split("_")
        out.add((int(p[0][1:]), int(p[1][1:]), int(p[2][2:]), int(p[3][4:]), int(p[4][1:]),
                 int(p[5][1:])))
    return sorted(out)


def lu_dat_key(M, N, NB, NRHS, P, Q) -> str:
    return f"M{M}_N{N}_NB{NB}_NRHS{NRHS}_P{P}_Q{Q}"


def c1() -> dict:
    z = np.load(HERE / "c1_n1024_nb64.npz")
    return {k: z[k] for k in z.files}


def c2() -> dict | None:
    p = HERE / "c2_n16384_nb256.npz"
    if not p.exists():
        return None
    z = np.load(p)
    return {k: z[k] for k in z.files}


def example_6x6():
    z = np.load(HERE / "example_6x6.npz")
    x = [Fraction(int(n), int(d)) for n, d in zip(z["x_num"], z["x_den"])]
    return np.asfortranarray(z["a"]), np.asfortranarray(z["b"]), x


def argcheck_cases() -> list[dict]:
    return json.loads((HERE / "argcheck_cases.json").read_text())


def aux_cases() -> list[tuple[int, int, int]]:
    """(N, NB, NRHS) of the stored cases of the routines beside the solve."""
    z = np.load(HERE / "aux_cases.npz")
    out = set()
    for k in z.files:
        p = k.split("_")
        out.add((int(p[0][1:]), int(p[1][2:]), int(p[2][4:])))
    return sorted(out)


def aux(N: int, NB: int, NRHS: int) -> dict:
    """Outputs of the reference's PDLANGE / PDGECON / PDGETRS('T') / PDGERFS
    for A = PDMATGEN(N, N, seed 100), B = PDMATGEN(N, NRHS, seed 200)."""
    z = np.load(HERE / "aux_cases.npz")
    pre = f"N{N}_NB{NB}_NRHS{NRHS}_"
    return {k[len(pre):]: z[k] for k in z.files if k.startswith(pre)}


def svx_cases() -> list[tuple[int, int, int]]:
    z = np.load(HERE / "svx_cases.npz")
    out = set()
    for k in z.files:
        p = k.split("_")
        out.add((int(p[0][1:]), int(p[1][2:]), int(p[2][4:])))
    return sorted(out)


def svx(N: int, NB: int, NRHS: int) -> dict:
    """Outputs of the reference's PDGEEQU / PDLAQGE / PDGESVX (see
    tests/golden/gen_golden.py)."""
    z = np.load(HERE / "svx_cases.npz")
    pre = f"N{N}_NB{NB}_NRHS{NRHS}_"
    return {k[len(pre):]: z[k] for k in z.files if k.startswith(pre)}


def svx_matrix(oracle, N: int) -> np.ndarray:
    """The scaled test matrix of the svx fixtures: PDMATGEN(N, N, 100) with
    A(i, j) * 2^(3 (i mod 7) - 4 (j mod 5)), 0-based indices."""
    a = oracle.matgen(N, N, 100)
    i = np.arange(N)[:, None] % 7
    j = np.arange(N)[None, :] % 5
    return np.asfortranarray(np.ldexp(a, 3 * i - 4 * j))
