"""The oracle (oracle/liboracle.so) pinned against the reference.

* the PDMATGEN LCG against an independent step-by-step restatement;
* IPIV / INFO / X against the golden fixtures the reference itself wrote
  (tests/golden/gen_golden.py: LU.dat cases, config 1, config 2);
* bit-for-bit against the reference build oracle/_ref when it is present;
* the residual checks against closed forms and planted mistakes;
* the 6 x 6 example against its exact rational solution.
"""
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
REF_LU = ROOT / "oracle" / "_ref" / "ref_lu"

A_MUL, C_ADD, MOD = 1103515245, 12345, 2 ** 31


def lcg_steps(seed, k):
    x = seed
    for _ in range(k):
        x = (A_MUL * x + C_ADD) % MOD
    return x


def test_lcg_jump_ahead_matches_stepping(oracle):
    for seed in (100, 200, 1, 12345):
        for k in (0, 1, 2, 3, 17, 1000, 4097):
            assert oracle.lcg_state(seed, k) == lcg_steps(seed, k)


def test_lcg_multiplier_is_the_reference_split():
    # pdmatgen.f:116-118: MULT0 = 20077, MULT1 = 16838 in base 2^16
    assert 16838 * 2 ** 16 + 20077 == A_MUL


def test_matgen_entries_follow_the_stream(oracle):
    M, N = 7, 5
    a = oracle.matgen(M, N, 100)
    x = 100
    for j in range(N):
        for i in range(M):
            x = (A_MUL * x + C_ADD) % MOD
            assert a[i, j] == 1.0 - 2.0 * (x / 2 ** 31)
    # a block of the global matrix equals the same entries of the whole
    blk = oracle.matgen(M, N, 100, i0=2, j0=1, mloc=4, nloc=3)
    assert np.array_equal(blk, a[2:6, 1:4])
    assert a.min() >= -1.0 and a.max() <= 1.0


@pytest.mark.parametrize("case_idx", range(0, 96, 1))
def test_oracle_matches_reference_lu_dat(oracle, golden, case_idx):
    cases = golden.lu_dat_cases()
    if case_idx >= len(cases):
        pytest.skip("no such case")
    M, N, NB, NRHS, P, Q = cases[case_idx]
    g = golden.lu_dat()
    key = golden.lu_dat_key(M, N, NB, NRHS, P, Q)
    a = oracle.matgen(M, N, 100)
    a0 = a.copy(order="F")
    ipiv, info = oracle.pdgetrf(a, NB)
    assert info == g[key + "_info"][0]
    # the reference's IPIV on any grid equals the 1 x 1 factorization's
    assert np.array_equal(ipiv, g[key + "_ipiv"]), (key, ipiv, g[key + "_ipiv"])
    assert oracle.fresid(a0, a, ipiv) < 1.0   # LU.dat threshold
    if M == N:
        b = oracle.matgen(N, NRHS, 200)
        x = oracle.pdgetrs(a, ipiv, b, NB)
        xr = g[key + "_x"]
        assert np.allclose(x, xr, rtol=1e-12, atol=1e-12 * np.abs(xr).max())
        assert oracle.sresid(a0, x, b) <= 1.0


def test_oracle_matches_reference_c1(oracle, golden):
    g = golden.c1()
    n = 1024
    a = oracle.matgen(n, n, 100)
    a0 = a.copy(order="F")
    b = np.asfortranarray(np.array([[sum(a0[i, :].tolist())] for i in range(n)]))
    ipiv, info, gap = oracle.pdgetrf(a, 64, with_gap=True)
    assert info == 0
    assert np.array_equal(ipiv, g["P1_Q1_ipiv"])
    assert np.array_equal(g["P1_Q1_ipiv"], g["P2_Q2_ipiv"])  # 2 x 2 grid, same pivots
    x = oracle.pdgetrs(a, ipiv, b, 64)
    assert np.allclose(x, g["P1_Q1_x"], rtol=1e-12, atol=1e-12)
    assert oracle.sresid(a0, x, b) < 1.0
    assert gap.min() > 0.0


def test_b_equals_a_times_ones_convention(oracle):
    """config 1's b = A * ones, summed left to right in double."""
    n = 64
    a = oracle.matgen(n, n, 100)
    s = np.zeros(n)
    for j in range(n):
        s = s + a[:, j]
    assert np.array_equal(s, np.array([sum(a[i, :].tolist()) for i in range(n)]))


def test_oracle_matches_reference_c2_pivots(oracle, golden):
    """config 2 (N = 16384): the reference's pivots are grid- and NB-robust;
    checked here on the leading 2048 x 2048 block's first 256 columns, where
    PDGETRF's first panel sees only A (no earlier updates)."""
    g = golden.c2()
    if g is None:
        pytest.skip("config 2 fixture not generated")
    n = 16384
    # The first panel's pivots depend only on A(:, 1:256).
    panel = oracle.matgen(n, 256, 100)
    piv, info = oracle.pdgetrf(panel, 256)
    assert info == 0
    assert np.array_equal(piv, g["ipiv"][:256])
    assert g["sresid"][0] < 1.0


@pytest.mark.skipif(not REF_LU.exists(), reason="reference build oracle/_ref absent")
@pytest.mark.parametrize("shape", [(300, 300, 32), (129, 77, 16), (64, 100, 8)])
def test_oracle_bit_identical_to_reference_build(oracle, shape):
    M, N, NB = shape
    env = dict(os.environ, MKL_THREADING_LAYER="SEQUENTIAL")
    with tempfile.TemporaryDirectory() as td:
        out = Path(td) / "o.bin"
        subprocess.run([str(REF_LU), str(M), str(N), str(NB), "1", "1", "1", "100", "lcg:200",
                        str(out), "--lu", "--nosolve"], check=True, env=env, capture_output=True)
        b = out.read_bytes()
        mn = min(M, N)
        ref_ipiv = np.frombuffer(b[72:72 + 4 * mn], np.int32)
        ref_lu = np.frombuffer(b[72 + 4 * mn:], np.float64).reshape(N, M).T
    a = oracle.matgen(M, N, 100)
    ipiv, info = oracle.pdgetrf(a, NB)
    assert np.array_equal(ipiv, ref_ipiv)
    assert np.array_equal(a, ref_lu)


def test_example_6x6_exact(oracle, golden):
    a, b, xs = golden.example_6x6()
    a0 = a.copy(order="F")
    ipiv, info = oracle.pdgetrf(a, 2)
    assert info == 0
    x = oracle.pdgetrs(a, ipiv, b, 2)
    exact = np.array([float(v) for v in xs])
    assert np.allclose(x[:, 0], exact, rtol=1e-14)
    assert oracle.sresid(a0, x, b) < 10.0   # EXAMPLE/pdscaex.f:188


def test_zero_pivot_info_and_completion(oracle):
    n = 6
    a = oracle.matgen(n, n, 7)
    a[:, 3] = 0.0                       # column 4 becomes exactly zero
    a[:, 4] = a[:, 1]                   # and column 5 dependent
    ipiv, info = oracle.pdgetrf(a, 2)
    assert info == 4                    # first exactly zero U(k,k), pdgetf2.f:226-228
    assert ipiv[3] == 4                 # PDAMAX keeps the row for a zero column
    b = np.asfortranarray([[1.0, 2.0], [2.0, 4.0]])
    ipiv, info = oracle.pdgetrf(b, 2)
    assert info == 2


def test_residual_checks_closed_forms(oracle):
    # diagonal A: LU = A, IPIV = identity, exact solution
    d = np.asfortranarray(np.diag([4.0, -2.0, 8.0, 1.0]))
    assert oracle.fresid(d, d, np.arange(1, 5, dtype=np.int32)) == 0.0
    b = np.asfortranarray([[4.0], [-2.0], [8.0], [1.0]])
    x = np.asfortranarray(np.ones((4, 1)))
    assert oracle.sresid(d, x, b) == 0.0
    # sresid scaling: r = e1 * eps*N*||A||*||x|| gives exactly 1
    eps = 2.0 ** -53
    b2 = b.copy(order="F")
    b2[0, 0] += 8.0 * 1.0 * eps * 4 * 2 ** 4   # representable perturbation
    assert oracle.sresid(d, x, b2) == pytest.approx(16.0)


def test_residual_checks_catch_planted_mistakes(oracle):
    n = 40
    a = oracle.matgen(n, n, 3)
    a0 = a.copy(order="F")
    ipiv, _ = oracle.pdgetrf(a, 8)
    assert oracle.fresid(a0, a, ipiv) < 1.0
    wrong_sign = a.copy(order="F")
    wrong_sign[5, 2] = -wrong_sign[5, 2]
    assert oracle.fresid(a0, wrong_sign, ipiv) > 1e6
    swapped = ipiv.copy()
    k = int(np.argmax(ipiv != np.arange(1, n + 1)))
    swapped[k] = k + 1
    assert oracle.fresid(a0, a, swapped) > 1e6
    b = oracle.matgen(n, 1, 9)
    x = oracle.pdgetrs(a, ipiv, b, 8)
    assert oracle.sresid(a0, x, b) < 1.0
    x[3, 0] += 1e-6
    assert oracle.sresid(a0, x, b) > 1e3


def test_pdlaswp_forward_backward_inverse(oracle):
    rng = np.random.default_rng(0)
    a = np.asfortranarray(rng.standard_normal((12, 5)))
    ipiv = np.array([3, 2, 7, 12, 5, 6, 9, 8, 12, 10, 11, 12], np.int32)
    f = a.copy(order="F")
    oracle.pdlaswp("F", f, 1, 5, 1, 8, ipiv)
    # brute force
    g = a.copy()
    for i in range(8):
        p = ipiv[i] - 1
        g[[i, p]] = g[[p, i]]
    assert np.array_equal(f, g)
    oracle.pdlaswp("B", f, 1, 5, 1, 8, ipiv)
    assert np.array_equal(f, a)
