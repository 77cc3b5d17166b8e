"""Host side: synthetic generator, reference-format table compiler, P2C and
total-degree start system, each pinned by independent numpy evaluation of the
polynomial system from its monomial description."""
import itertools

import numpy as np
import pytest


def np_eval(exps, eq, coeffs, x):
    """H, J from the monomial description (independent of any table)."""
    Cn, n = exps.shape
    H = np.zeros(n, dtype=complex)
    J = np.zeros((n, n), dtype=complex)
    for j in range(Cn):
        m = np.prod(x ** exps[j])
        H[eq[j]] += coeffs[j] * m
        for v in range(n):
            e = exps[j, v]
            if e:
                ee = exps[j].copy()
                ee[v] -= 1
                J[eq[j], v] += coeffs[j] * e * np.prod(x ** ee)
    return H, J


def test_synthetic_supports(hcx_mod):
    s = hcx_mod.System.synthetic(6, 24, 6, 12, 42)
    exps, eq = s.monomials()
    assert s.info["num_coeffs"] == 72
    for i in range(6):
        rows = exps[eq == i]
        assert len(rows) == 12
        assert any((r == 0).all() for r in rows)  # constant
        for v in range(6):  # all pure squares
            e = np.zeros(6, int)
            e[v] = 2
            assert any((r == e).all() for r in rows)
        assert rows.sum(axis=1).max() == 2
        assert len({tuple(r) for r in rows}) == 12  # distinct
    # mixed degrees: 6 quadratic + 6 cubic, 20 monomials each
    s4 = hcx_mod.System.synthetic(12, 36, 6, 20, 4)
    e4, q4 = s4.monomials()
    for i in range(12):
        d = 2 if i < 6 else 3
        rows = e4[q4 == i]
        assert len(rows) == 20 and rows.sum(axis=1).max() == d
        for v in range(12):
            assert any(r[v] == d and r.sum() == d for r in rows)
    # dense n=3 quadratic: all 10 monomials of degree <= 2
    s1 = hcx_mod.System.synthetic(3, 0, 3, 0, 1)
    e1, q1 = s1.monomials()
    assert s1.info["num_coeffs"] == 30 and all((q1 == i).sum() == 10 for i in range(3))


def test_synthetic_determinism(hcx_mod):
    a = hcx_mod.System.synthetic(6, 24, 6, 12, 5)
    b = hcx_mod.System.synthetic(6, 24, 6, 12, 5)
    c = hcx_mod.System.synthetic(6, 24, 6, 12, 6)
    assert (a.monomials()[0] == b.monomials()[0]).all()
    assert (a.affine() == b.affine()).all()
    assert not (a.affine() == c.affine()).all()


def test_random_streams(hcx_mod):
    z = hcx_mod.random_cn(123, 200000)
    assert abs(z.mean()) < 0.01
    assert abs(np.mean(np.abs(z) ** 2) - 1.0) < 0.01
    assert abs(np.mean(z.real ** 2) - 0.5) < 0.01 and abs(np.mean(z.real * z.imag)) < 0.01
    assert (hcx_mod.random_cn(123, 10) == z[:10]).all()
    g = hcx_mod.random_gamma(9)
    assert abs(abs(g) - 1) < 1e-15


@pytest.mark.parametrize("n,k,nq,mono,seed", [(3, 0, 3, 0, 1), (6, 24, 6, 12, 2), (5, 4, 2, 9, 3),
                                               (12, 36, 6, 20, 4)])
def test_tables_match_polynomial(hcx_mod, oracle_mod, n, k, nq, mono, seed):
    """Reference-format tables + the oracle's table evaluator reproduce H, dH/dx
    and dH/dt computed directly from the monomials."""
    s = hcx_mod.System.synthetic(n, k, nq, mono, seed)
    exps, eq = s.monomials()
    hx_idx, ht_idx = s.tables()
    Cn = s.info["num_coeffs"]
    rng = np.random.default_rng(seed)
    c0 = rng.normal(size=Cn) + 1j * rng.normal(size=Cn)
    c1 = rng.normal(size=Cn) + 1j * rng.normal(size=Cn)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.7
    t = 0.37
    hx = np.zeros((Cn + 1, 2), complex)
    hx[:Cn, 0], hx[:Cn, 1], hx[Cn, 0] = c0, c1, 1.0
    ht = np.zeros((Cn + 1, 1), complex)
    ht[:Cn, 0] = c1
    op = oracle_mod.OracleProblem(n, Cn, hx_idx, ht_idx, 1)
    H, J, Ht = op.eval(hx, ht, x, t)
    Href, Jref = np_eval(exps, eq, c0 + t * c1, x)
    Htref, _ = np_eval(exps, eq, c1, x)
    scale = 1 + np.abs(Href).max()
    assert np.abs(H - Href).max() < 1e-12 * scale
    assert np.abs(J - Jref).max() < 1e-12 * (1 + np.abs(Jref).max())
    assert np.abs(Ht - Htref).max() < 1e-12 * (1 + np.abs(Htref).max())
    # finite differences pin the Jacobian and dH/dt independently
    h = 1e-6
    for v in range(min(n, 3)):
        e = np.zeros(n, complex)
        e[v] = h
        fd = (op.eval(hx, ht, x + e, t)[0] - op.eval(hx, ht, x - e, t)[0]) / (2 * h)
        assert np.abs(fd - J[:, v]).max() < 1e-6 * (1 + np.abs(J).max())
    fdt = (op.eval(hx, ht, x, t + h)[0] - op.eval(hx, ht, x, t - h)[0]) / (2 * h)
    assert np.abs(fdt - Ht).max() < 1e-6 * (1 + np.abs(Ht).max())


def test_table_padding_layout(hcx_mod):
    """Padding terms are {0, C, n, ...}; derivative scalars are exponents."""
    s = hcx_mod.System.synthetic(4, 0, 0, 0, 3)  # dense cubic
    hx_idx, ht_idx = s.tables()
    Cn, n = s.info["num_coeffs"], 4
    pad = hx_idx[hx_idx[..., 0] == 0]
    assert (pad[:, 1] == Cn).all() and (pad[:, 2:] == n).all()
    assert set(np.unique(hx_idx[..., 0])) <= {0, 1, 2, 3}
    assert s.info["hx_parts"] == 4 and s.info["ht_parts"] == 5


def test_p2c_host_matches_oracle_and_numpy(hcx_mod, oracle_mod):
    s = hcx_mod.System.synthetic(6, 24, 6, 12, 8)
    A = s.affine()
    p0 = hcx_mod.random_cn(1, 24)
    p1 = hcx_mod.random_cn(2, 24)
    hx, ht = s.param_homotopy_coeffs(p0, p1)
    c0 = A @ np.concatenate([[1], p0])
    c1 = A @ np.concatenate([[1], p1])
    assert np.allclose(hx[:-1, 0], c0, rtol=0, atol=1e-12)
    assert np.allclose(hx[:-1, 1], c1 - c0, rtol=0, atol=1e-12)
    assert np.allclose(ht[:-1, 0], c1 - c0, rtol=0, atol=1e-12)
    assert hx[-1, 0] == 1 and hx[-1, 1] == 0 and ht[-1, 0] == 0
    ohx, oht = oracle_mod.p2c_affine(A, p0, p1[None])
    assert np.abs(ohx[0] - hx).max() < 1e-12 and np.abs(oht[0] - ht).max() < 1e-12
    assert np.allclose(s.eval_coeffs(p0), c0, atol=1e-12)


def test_total_degree_homotopy(hcx_mod, oracle_mod):
    s = hcx_mod.System.synthetic(4, 0, 2, 8, 11)  # 2 quadratic + 2 cubic
    f = s.eval_coeffs(None)
    gamma = hcx_mod.random_gamma(3)
    td, hx, ht = s.total_degree(f, gamma)
    inf = td.info
    n, C2 = 4, inf["num_coeffs"]
    idx = np.arange(2 * 2 * 3 * 3)
    roots = td.start_roots(idx)
    degs = [2, 2, 3, 3]
    for v in range(n):
        assert np.allclose(roots[:, v] ** degs[v], 1, atol=1e-14)
    assert len({tuple(np.round(r, 9)) for r in roots}) == 36
    hxi, hti = td.tables()
    op = oracle_mod.OracleProblem(n, C2, hxi, hti, 1)
    e2, q2 = td.monomials()
    for r in roots[:5]:
        H0 = op.eval(hx, ht, r, 0.0)[0]
        assert np.abs(H0).max() < 1e-13
    # t = 1 is the target system F
    x = np.array([0.3 + 0.1j, -0.7j, 1.1, 0.2 - 0.4j])
    H1 = op.eval(hx, ht, x, 1.0)[0]
    e1, q1 = s.monomials()
    Hf, _ = np_eval(e1, q1, f, x)
    assert np.abs(H1 - Hf).max() < 1e-12
    # t = 0 is gamma * (x_i^d - 1)
    G = np.array([x[i] ** degs[i] - 1 for i in range(n)]) * gamma
    assert np.abs(op.eval(hx, ht, x, 0.0)[0] - G).max() < 1e-12


def test_total_degree_subset(hcx_mod):
    s = hcx_mod.System.synthetic(9, 0, 9, 15, 1)
    idx = s.subset(142, 5)
    assert len(set(idx.tolist())) == 142 and idx.min() >= 0 and idx.max() < 2 ** 9
    assert (s.subset(142, 5) == idx).all()
    assert (hcx_mod.System.synthetic(3, 0, 3, 0, 1).subset(8, 1) == np.arange(8)).all()
    big = hcx_mod.System.synthetic(30, 0, 30, 36, 1).subset(312, 9)
    assert len(set(big.tolist())) == 312 and big.max() < 2 ** 30


def test_collect_host_semantics(hcx_mod, oracle_mod):
    rng = np.random.default_rng(0)
    n, I, K = 3, 2, 6
    roots = rng.normal(size=(I * K, n)) + 1j * rng.normal(size=(I * K, n))
    roots[3] = roots[1] + 1e-6  # duplicate of path 1
    roots[9] = roots[7]
    flags = np.full(I * K, hcx_mod.FLAG_TRACKED | hcx_mod.FLAG_REFINED, np.uint32)
    flags[5] = hcx_mod.FLAG_INFINITY
    flags[0] |= hcx_mod.FLAG_REAL
    counts, f2 = hcx_mod.collect_host(n, I, K, roots, flags)
    oc, of = oracle_mod.collect(n, I, K, roots, flags)
    assert counts == oc and (f2 == of).all()
    assert counts["refined"] == 11 and counts["unique"] == 9 and counts["infinity"] == 1 and counts["real"] == 1
    assert not (f2[3] & hcx_mod.FLAG_UNIQUE) and (f2[1] & hcx_mod.FLAG_UNIQUE)
