"""Pins the oracle (test infrastructure) against what mathematics fixes:
LAPACK's zgetrf pivot sequence, numpy solves, exact tracking of linear and
polynomial solution paths, and complete Bezout root sets."""
import numpy as np
import pytest
import scipy.linalg


@pytest.mark.parametrize("n", [1, 2, 3, 6, 12, 17, 32])
def test_lu_matches_lapack(oracle_mod, n):
    rng = np.random.default_rng(n)
    for _ in range(5):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        b = rng.normal(size=n) + 1j * rng.normal(size=n)
        x, piv, info = oracle_mod.lu_solve(A, b)
        assert info == 0
        xr = np.linalg.solve(A, b)
        assert np.abs(x - xr).max() <= 1e-11 * np.linalg.cond(A) * (1 + np.abs(xr).max())
        # zgetrf pivots with izamax = first max of |Re|+|Im|, the same rule
        _, lpiv = scipy.linalg.lu_factor(A)
        assert (piv == lpiv).all()


def test_lu_ties_and_zero_columns(oracle_mod):
    # equal cabs1 candidates: the first (in current row order) wins
    A = np.array([[1, 2, 0], [1j, 1, 1], [-1, 0, 3]], dtype=complex)
    x, piv, info = oracle_mod.lu_solve(A, np.ones(3))
    assert piv[0] == 0 and info == 0
    A2 = np.array([[0.5, 1], [1 + 0j, 2], ], dtype=complex)
    _, piv2, _ = oracle_mod.lu_solve(A2, np.ones(2))
    assert piv2[0] == 1
    # zero column -> info = k+1, elimination skipped
    Z = np.array([[1, 0, 2], [3, 0, 1], [1, 0, 1]], dtype=complex)
    _, _, info = oracle_mod.lu_solve(Z, np.ones(3))
    assert info == 2


def linear_problem(hcx_mod, oracle_mod, a, b):
    """H(x,t) = x - ((1-t) a + t b): monomials x_i and 1 per equation."""
    n = len(a)
    exps, eq = [], []
    for i in range(n):
        e = np.zeros(n, int)
        e[i] = 1
        exps += [e, np.zeros(n, int)]
        eq += [i, i]
    s = hcx_mod.System.from_monomials(np.array(exps), np.array(eq))
    hx_idx, ht_idx = s.tables()
    Cn = 2 * n
    hx = np.zeros((Cn + 1, 2), complex)
    ht = np.zeros((Cn + 1, 1), complex)
    for i in range(n):
        hx[2 * i] = [1, 0]
        hx[2 * i + 1] = [-a[i], -(b[i] - a[i])]
        ht[2 * i + 1, 0] = -(b[i] - a[i])
    hx[Cn, 0] = 1
    return s, oracle_mod.OracleProblem(n, Cn, hx_idx, ht_idx, 1), hx, ht


def test_linear_homotopy_exact(hcx_mod, oracle_mod):
    a = np.array([1 + 2j, -0.5j, 3.0])
    b = np.array([-2 + 1j, 0.25, 1j])
    _, op, hx, ht = linear_problem(hcx_mod, oracle_mod, a, b)
    out = op.track(a[None], hx[None], ht[None], threads=1, dt_min=0.0)
    assert out["flags"][0] & oracle_mod.FLAG_TRACKED and out["flags"][0] & oracle_mod.FLAG_REFINED
    assert np.abs(out["roots"][0] - b).max() < 1e-13
    assert out["newton"][0] == out["steps"][0]  # every corrector succeeds at once
    assert abs(out["t_final"][0] - 1) <= 1e-7


def test_rk4_exact_for_quartic_paths(hcx_mod, oracle_mod):
    """x(t) = g(t) with deg g = 4: RK4 (Simpson) integrates x' = g'(t) exactly,
    so with a near-zero corrector tolerance every step is still accepted."""
    n = 2
    g = np.array([[0.3 + 1j, 1.0, -2j, 0.5, 0.25 + 0.5j], [-1.0, 0.2j, 0.7, -0.3j, 1.1]])  # [var][power]
    exps, eq = [], []
    for i in range(n):
        e = np.zeros(n, int)
        e[i] = 1
        exps += [e, np.zeros(n, int)]
        eq += [i, i]
    s = hcx_mod.System.from_monomials(np.array(exps), np.array(eq))
    hx_idx, ht_idx = s.tables()
    M, Cn = 4, 2 * n
    hx = np.zeros((Cn + 1, M + 1), complex)
    ht = np.zeros((Cn + 1, M), complex)
    for i in range(n):
        hx[2 * i, 0] = 1
        hx[2 * i + 1] = -g[i]
        ht[2 * i + 1] = -g[i, 1:] * np.arange(1, M + 1)
    hx[Cn, 0] = 1
    op = oracle_mod.OracleProblem(n, Cn, hx_idx, ht_idx, M)
    out = op.track(g[:, 0][None], hx[None], ht[None], threads=1, corr_tol=1e-26, dt_min=0.0)
    assert out["flags"][0] & oracle_mod.FLAG_TRACKED
    assert out["newton"][0] == out["steps"][0]  # no rejection at all
    assert np.abs(out["roots"][0] - g.sum(axis=1)).max() < 1e-13
    # control: classic forward Euler would need rejections under the same tolerance;
    # a wrongly weighted predictor shows up as rejected steps
    assert out["steps"][0] < 40


def config1_oracle(hcx_mod, oracle_mod, seed):
    from hcx import workloads
    c = workloads.config1(seed)
    td = c["homotopy"]
    hxi, hti = td.tables()
    op = oracle_mod.OracleProblem(3, td.info["num_coeffs"], hxi, hti, 1)
    return c, op, op.track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_config1_finds_all_bezout_roots(hcx_mod, oracle_mod, seed):
    """A generic dense quadratic 3x3 system has exactly 8 finite roots; 8
    distinct nonsingular roots with zero residual are therefore all of them."""
    from test_system import np_eval
    c, op, out = config1_oracle(hcx_mod, oracle_mod, seed)
    assert ((out["flags"] & oracle_mod.FLAG_REFINED) != 0).all()
    sysf = c["system"]
    exps, eq = sysf.monomials()
    f = sysf.eval_coeffs(None)
    for r in out["roots"]:
        H, J = np_eval(exps, eq, f, r)
        assert np.abs(H).max() < 1e-12 * (1 + np.abs(r).max() ** 2)
        assert np.linalg.cond(J) < 1e10
    counts, _ = oracle_mod.collect(3, 1, 8, out["roots"], out["flags"])
    assert counts["unique"] == 8
    assert out["residual"].max() < 1e-11


def test_failure_paths(hcx_mod, oracle_mod):
    c, op, _ = config1_oracle(hcx_mod, oracle_mod, 1)
    out = op.track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1, max_steps=1)
    assert (out["flags"] & oracle_mod.FLAG_BUDGET).all()
    assert not (out["flags"] & oracle_mod.FLAG_TRACKED).any()
    out = op.track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1, inf_tol=1e-3)
    assert (out["flags"] & oracle_mod.FLAG_INFINITY).all()


def test_threads_do_not_change_results(hcx_mod, oracle_mod):
    c, op, out1 = config1_oracle(hcx_mod, oracle_mod, 2)
    hx = np.repeat(c["coeffs_hx"], 5, axis=0)
    ht = np.repeat(c["coeffs_ht"], 5, axis=0)
    outm = op.track(c["start"], hx, ht, threads=4)
    assert (outm["roots"][:8] == out1["roots"]).all()
    assert (outm["roots"][8:16] == out1["roots"]).all()
