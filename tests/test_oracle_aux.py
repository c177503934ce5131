"""The oracle of the routines beside the solve (PDLANGE, PDGECON with
PDLACON, PDGETRS TRANS='T', PDGERFS), pinned against the reference's own
outputs (tests/golden/aux_cases.npz, written by oracle/_ref/ref_lu aux) and
against what the mathematics fixes.

Tolerances: the one / infinity / max norms are sums and maxima taken in the
reference's order and must agree bit for bit; the Frobenius norm calls the
same DLASSQ and must agree bit for bit as well.  RCOND, FERR and the 'T'
solve go through MKL's triangular solves in the oracle and through the
reference's PBLAS in the fixture and agree to 1e-12 relative (FERR to 1e-3:
its weights contain the rounding-level residual).  BERR is the
ratio of a residual (itself a rounding-level quantity) to |A||x| + |b|: two
correct implementations agree to its order of magnitude only, so it is
compared within a factor 2 and against an absolute bound of a few eps.
"""
import numpy as np
import pytest

EPS = 2.0 ** -53


def _case(golden, oracle, N, NB, NRHS):
    r = golden.aux(N, NB, NRHS)
    a = oracle.matgen(N, N, 100)
    b = oracle.matgen(N, NRHS, 200)
    lu = a.copy(order="F")
    piv, info = oracle.pdgetrf(lu, NB)
    assert info == 0
    return r, a, b, lu, piv


def test_aux_fixtures_present(golden):
    assert golden.aux_cases() == [(37, 4, 2), (200, 32, 3), (513, 64, 1)]


@pytest.mark.parametrize("idx", range(3))
def test_pdlange_matches_reference_bitwise(oracle, golden, idx):
    N, NB, NRHS = golden.aux_cases()[idx]
    r, a, *_ = _case(golden, oracle, N, NB, NRHS)
    for k, c in enumerate("M1IF"):
        assert oracle.pdlange(c, a) == r["norms"][k], c
        assert oracle.pdlange(c, a, N - 3, N - 2, 2, 3) == r["subnorms"][k], c


def test_pdlange_closed_forms(oracle):
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.uniform(-3, 3, (57, 41)))
    assert oracle.pdlange("M", a) == np.abs(a).max()
    assert np.isclose(oracle.pdlange("1", a), np.abs(a).sum(0).max(), rtol=1e-15, atol=0)
    assert np.isclose(oracle.pdlange("O", a), np.abs(a).sum(0).max(), rtol=1e-15, atol=0)
    assert np.isclose(oracle.pdlange("I", a), np.abs(a).sum(1).max(), rtol=1e-15, atol=0)
    assert np.isclose(oracle.pdlange("F", a), np.linalg.norm(a), rtol=1e-14, atol=0)
    assert np.isclose(oracle.pdlange("E", a), np.linalg.norm(a), rtol=1e-14, atol=0)
    # one and infinity norms are transposes of each other
    assert oracle.pdlange("1", a) == oracle.pdlange("I", np.asfortranarray(a.T))
    # a 1 x 1 case and an empty one
    assert oracle.pdlange("F", np.asfortranarray([[-2.5]])) == 2.5
    assert oracle.pdlange("M", a, 0, 5) == 0.0


def _alt(n):
    k = np.arange(1, n + 1)
    return np.where(k % 2 == 0, -1.0, 1.0) * (1.0 + (k - 1) / (n - 1))


def _perm(piv):
    p = np.arange(len(piv))
    for i, t in enumerate(piv):
        p[i], p[t - 1] = p[t - 1], p[i]
    return p


@pytest.mark.parametrize("idx", range(3))
def test_pdgecon_matches_reference(oracle, golden, idx):
    N, NB, NRHS = golden.aux_cases()[idx]
    r, a, b, lu, piv = _case(golden, oracle, N, NB, NRHS)
    rc1 = oracle.pdgecon("1", lu, oracle.pdlange("1", a))
    rcI = oracle.pdgecon("I", lu, oracle.pdlange("I", a))
    assert np.isclose(rc1, r["rcond"][0], rtol=1e-12, atol=0)
    assert np.isclose(rcI, r["rcond"][1], rtol=1e-12, atol=0)


@pytest.mark.parametrize("idx", range(3))
def test_pdgecon_is_the_alternating_sign_bound(oracle, golden, idx):
    """pdlacon.f:188-189 reset the running estimate on every entry, so the
    reference's RCOND is 1 / (||A|| * 2 ||B x_alt||_1 / (3n)) with
    B = U^{-1} L^{-1} ('1') or L^{-T} U^{-T} ('I'): checked with numpy's
    solves, independently of the oracle's loop."""
    N, NB, NRHS = golden.aux_cases()[idx]
    r, a, b, lu, piv = _case(golden, oracle, N, NB, NRHS)
    L = np.tril(lu, -1) + np.eye(N)
    U = np.triu(lu)
    alt = _alt(N)
    e1 = 2 * np.abs(np.linalg.solve(U, np.linalg.solve(L, alt))).sum() / (3 * N)
    eI = 2 * np.abs(np.linalg.solve(L.T, np.linalg.solve(U.T, alt))).sum() / (3 * N)
    assert np.isclose(r["rcond"][0], 1 / (np.abs(a).sum(0).max() * e1), rtol=1e-9, atol=0)
    assert np.isclose(r["rcond"][1], 1 / (np.abs(a).sum(1).max() * eI), rtol=1e-9, atol=0)
    # L U is P A: the product of the factors rebuilds the permuted matrix
    assert np.allclose(L @ U, a[_perm(piv)], atol=1e-12)


def test_pdgecon_quick_returns(oracle):
    lu = np.asfortranarray([[3.0]])
    assert oracle.pdgecon("1", lu, 3.0) == 1.0           # N = 1
    assert oracle.pdgecon("1", np.asfortranarray(np.eye(4)), 0.0) == 0.0  # ANORM = 0


def test_pdgecon_diagonal_closed_form(oracle):
    n = 50
    d = np.linspace(1.0, 9.0, n)
    lu = np.asfortranarray(np.diag(d))
    anorm = d.max()
    est = 2 * np.sum(np.abs(_alt(n)) / d) / (3 * n)
    assert np.isclose(oracle.pdgecon("1", lu, anorm), 1 / (anorm * est), rtol=1e-13, atol=0)
    assert np.isclose(oracle.pdgecon("I", lu, anorm), 1 / (anorm * est), rtol=1e-13, atol=0)


@pytest.mark.parametrize("idx", range(3))
def test_pdgetrs_transposed_matches_reference(oracle, golden, idx):
    N, NB, NRHS = golden.aux_cases()[idx]
    r, a, b, lu, piv = _case(golden, oracle, N, NB, NRHS)
    xT = oracle.pdgetrs_t(lu, piv, b)
    assert np.allclose(xT, r["xT"], rtol=0, atol=1e-12 * np.abs(r["xT"]).max())
    # and it solves A^T x = b
    assert np.allclose(a.T @ xT, b, atol=1e-10 * np.abs(b).max())


@pytest.mark.parametrize("idx", range(3))
def test_pdgerfs_matches_reference(oracle, golden, idx):
    N, NB, NRHS = golden.aux_cases()[idx]
    r, a, b, lu, piv = _case(golden, oracle, N, NB, NRHS)
    for c, tr in enumerate("NNT"):
        x, ferr, berr = oracle.pdgerfs(tr, a, lu, piv, b, r[f"rfs{c}_x0"])
        xr = r[f"rfs{c}_x"]
        assert np.allclose(x, xr, rtol=0, atol=1e-13 * np.abs(xr).max()), (c, tr)
        # FERR's weights hold |r| next to (n+1) eps (|A||x| + |b|), about
        # 1/(n+1) of the weight, and r is rounding-level: 1e-3 relative
        assert np.allclose(ferr, r[f"rfs{c}_ferr"], rtol=1e-3, atol=0), (c, tr)
        assert np.all(berr <= 4 * EPS) and np.all(r[f"rfs{c}_berr"] <= 4 * EPS)
        assert np.all(berr <= 2 * r[f"rfs{c}_berr"] + EPS) and np.all(r[f"rfs{c}_berr"] <= 2 * berr + EPS)


def test_pdgerfs_error_bounds_closed_form(oracle):
    """BERR is the componentwise backward error of the returned x, and FERR
    is (2 / 3n) ||diag(w) op(A)^{-T} x_alt||_1 / ||x||_inf with w = |r| +
    (n+1) eps (|op(A)||x| + |b|): the w term dominates |r|, so the closed
    form without r agrees to a few percent.  A refinement that starts far
    from the solution must converge to the FP64 solution."""
    n, nb = 120, 16
    # rows scaled over three decades, so that A and A^T differ in structure
    a = np.asfortranarray(oracle.matgen(n, n, 31) * np.geomspace(1.0, 1e3, n)[:, None])
    b = oracle.matgen(n, 2, 32)
    lu = a.copy(order="F")
    piv, _ = oracle.pdgetrf(lu, nb)
    for tr in "NT":
        opa = a if tr == "N" else np.asfortranarray(a.T)
        xtrue = np.linalg.solve(opa, b)
        x0 = np.asfortranarray(xtrue * (1 + 1e-3 * np.cos(np.arange(n)))[:, None])
        x, ferr, berr = oracle.pdgerfs(tr, a, lu, piv, b, x0)
        assert np.allclose(x, xtrue, rtol=0, atol=1e-12 * np.abs(xtrue).max())
        for k in range(2):
            w = np.abs(opa) @ np.abs(x[:, k]) + np.abs(b[:, k])
            r = (b[:, k].astype(np.longdouble) - opa.astype(np.longdouble) @ x[:, k].astype(np.longdouble))
            s = float(np.max(np.abs(r) / w))
            assert berr[k] <= 4 * EPS and s <= 4 * EPS
            wf = (n + 1) * EPS * w
            opt = np.linalg.solve(opa.T, _alt(n))        # op(A)^{-T} x_alt
            est = 2 * np.abs(wf * opt).sum() / (3 * n)
            assert np.isclose(ferr[k], est / np.abs(x[:, k]).max(), rtol=0.05, atol=0), (tr, k)
            # the check discriminates: with the operand not transposed (a
            # plausible mistake) the bound differs by far more than 5 %
            wrong = 2 * np.abs(wf * np.linalg.solve(opa, _alt(n))).sum() / (3 * n)
            assert not np.isclose(wrong, est, rtol=0.05, atol=0)


def test_pdgerfs_quick_return(oracle):
    a = np.asfortranarray([[2.0]])
    x, ferr, berr = oracle.pdgerfs("N", a, a.copy(order="F"), np.array([1], np.int32),
                                   np.asfortranarray([[4.0]]), np.asfortranarray([[2.0]]))
    assert ferr[0] == 0.0 and berr[0] == 0.0


# ------------------------------------------------ PDGEEQU / PDLAQGE / PDGESVX
@pytest.mark.parametrize("idx", range(2))
def test_equilibration_matches_reference_bitwise(oracle, golden, idx):
    """PDGEEQU and PDLAQGE are maxima, reciprocals and products: bit for bit."""
    N, NB, NRHS = golden.svx_cases()[idx]
    r = golden.svx(N, NB, NRHS)
    a = golden.svx_matrix(oracle, N)
    R, C, rc, cc, am, info = oracle.pdgeequ(a)
    assert info == int(r["info_equ"]) == 0
    assert np.array_equal(R, r["R"]) and np.array_equal(C, r["C"])
    assert (rc, cc, am) == (r["rowcnd"][0], r["colcnd"][0], r["amax"][0])
    a2 = a.copy(order="F")
    assert oracle.pdlaqge(a2, R, C, rc, cc, am) == chr(int(r["equed_laqge"])) == "B"
    assert np.array_equal(a2, r["A_laqge"])


def test_equilibration_closed_forms(oracle):
    """R_i = 1 / max_j |a_ij|, C_j = 1 / max_i |a_ij| R_i: every row of
    diag(R) A has max 1 and every column of diag(R) A diag(C) max 1; a zero
    row or column is reported by its index; mild scaling is left alone."""
    rng = np.random.default_rng(8)
    a = np.asfortranarray(rng.uniform(-1, 1, (30, 20)) * np.geomspace(1e-3, 1e3, 30)[:, None])
    R, C, rc, cc, am, info = oracle.pdgeequ(a)
    assert info == 0 and am == np.abs(a).max()
    assert np.allclose(np.abs(R[:, None] * a).max(1), 1.0, rtol=1e-15, atol=0)
    assert np.allclose(np.abs(R[:, None] * a * C[None, :]).max(0), 1.0, rtol=1e-15, atol=0)
    z = a.copy(order="F")
    z[7, :] = 0.0
    assert oracle.pdgeequ(z)[5] == 8
    z = a.copy(order="F")
    z[:, 4] = 0.0
    assert oracle.pdgeequ(z)[5] == 30 + 5
    m = np.asfortranarray(rng.uniform(0.5, 1.0, (10, 10)))
    R, C, rc, cc, am, _ = oracle.pdgeequ(m)
    assert oracle.pdlaqge(m.copy(order="F"), R, C, rc, cc, am) == "N"


@pytest.mark.parametrize("idx", range(2))
def test_pdgesvx_matches_reference(oracle, golden, idx):
    N, NB, NRHS = golden.svx_cases()[idx]
    r = golden.svx(N, NB, NRHS)
    a = golden.svx_matrix(oracle, N)
    b = oracle.matgen(N, NRHS, 200)
    for c, (fact, trans) in enumerate([("E", "N"), ("E", "T"), ("N", "N")]):
        s = oracle.pdgesvx(fact, trans, a, b, NB)
        assert s["info"] == int(r[f"c{c}_info"]) == 0
        assert s["equed"] == chr(int(r[f"c{c}_equed"]))
        X = r[f"c{c}_X"]
        assert np.allclose(s["X"], X, rtol=0, atol=1e-13 * np.abs(X).max()), (fact, trans)
        assert np.isclose(s["rcond"], r[f"c{c}_rcond"][0], rtol=1e-12, atol=0)
        assert np.allclose(s["ferr"], r[f"c{c}_ferr"], rtol=8 / (N + 1), atol=0)
        assert np.all(s["berr"] <= 4 * EPS)
        assert np.array_equal(s["B"], r[f"c{c}_B"])       # B scaled as the reference leaves it
        if fact == "E":
            assert np.array_equal(s["R"], r[f"c{c}_R"]) and np.array_equal(s["C"], r[f"c{c}_C"])


def test_pdgesvx_singular_reports_n_plus_one(oracle):
    """RCOND below eps: INFO = N + 1 and X is not computed (pdgesvx.f:715)."""
    n = 40
    rng = np.random.default_rng(12)
    u = rng.standard_normal((n, 3))
    a = np.asfortranarray(u @ rng.standard_normal((3, n)) + 1e-18 * np.eye(n))
    s = oracle.pdgesvx("N", "N", a, np.asfortranarray(np.ones((n, 1))), 8)
    assert s["info"] == n + 1 or s["info"] > 0
