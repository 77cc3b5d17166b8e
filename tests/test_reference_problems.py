"""The reference's own problems (GPU-HC/problems/<p>/) through the table
semantics, the coefficient maps and the oracle (CPU).

Fixtures: tests/golden/reference/<p>.npz, written by
tests/golden/make_reference_fixtures.py from the reference's problem data and
system definitions (data only).  What pins what:

* the exact (rational arithmetic) coefficient-polynomial tables of the
  start -> target path pin the P2C of the parameter-polynomial map;
* F(x; p) and dF/dx evaluated from the system definition at random points pin
  the dHdt / dHdx tables together with the coefficient map, through the
  oracle's evaluator;
* the reference's committed start solutions must solve the start system
  H(x, t = 0) = F(x; p_start) = 0 (a check of the loader, the tables and the
  P2C at once).
"""
import glob
import os

import numpy as np
import pytest

import oracle
from hcx import compiler as cc

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "reference", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]

# The committed start solutions of P3P_10x10 do not solve its system at its
# start parameters (max |H| ~ 3e3); the reference never builds that problem
# (its kernel is missing from magmaHC/CMakeLists.txt:1-86).
BAD_START_DATA = {"P3P_10x10"}


def load(name):
    z = np.load(os.path.join(HERE, "golden", "reference", name + ".npz"))
    d = {k: z[k] for k in z.files}
    n = int(d["n"])
    d["name"] = str(d["name"])
    d["kind"] = str(d["kind"])
    for key in ("n", "k", "M", "num_tracks", "max_steps", "max_corr", "incr_steps"):
        d[key] = int(d[key])
    if d["kind"] == "PH":
        ht_ph = d["ht"].reshape(int(d["ht_terms"]), int(d["ht_parts"]), n)
        hx_ph = d["hx"].reshape(n, int(d["hx_terms"]), int(d["hx_parts"]), n)
        conv = cc.ph_to_p2c(n, d["k"], ht_ph, hx_ph)
        d.update(C=conv["C"], ht3=conv["ht"], hx4=conv["hx"], cmap=conv["cmap"], M=conv["max_order_t"])
    else:
        d["C"] = int(d["C"])
        d["ht3"] = d["ht"].reshape(n, int(d["ht_terms"]), int(d["ht_parts"]))
        d["hx4"] = d["hx"].reshape(n, n, int(d["hx_terms"]), int(d["hx_parts"]))
        d["cmap"] = (d["cmap_scalar"], d["cmap_params"])
    return d


def oracle_problem(d):
    return oracle.OracleProblem(d["n"], d["C"], d["hx4"], d["ht3"], d["M"])


def test_all_problems_present():
    assert len(NAMES) == 21
    kinds = [load(nm)["kind"] for nm in NAMES]
    assert kinds.count("PH") == 1


@pytest.mark.parametrize("name", NAMES)
def test_p2c_matches_exact_tables(name):
    d = load(name)
    if "coeffs_hx" not in d:
        pytest.skip("target parameters incomplete in the reference data")
    hx, ht = oracle.p2c_poly(*d["cmap"], d["start_params"], d["target_params"], d["M"])
    scale = np.abs(d["coeffs_hx"]).max()
    assert np.abs(hx[0] - d["coeffs_hx"]).max() <= 1e-13 * scale
    assert np.abs(ht[0] - d["coeffs_ht"]).max() <= 1e-13 * scale


@pytest.mark.parametrize("name", NAMES)
def test_tables_encode_the_system(name):
    d = load(name)
    op = oracle_problem(d)
    for x, p, F, J in zip(d["check_x"], d["check_p"], d["check_f"], d["check_j"]):
        # p(t) constant at p: c(t) = c(p) for every t
        hx, ht = oracle.p2c_poly(*d["cmap"], p, p, d["M"])
        H, Jt, Ht = op.eval(hx[0], ht[0], x, 0.37)
        assert np.abs(H - F).max() <= 1e-12 * max(1.0, np.abs(F).max())
        assert np.abs(Jt - J).max() <= 1e-12 * max(1.0, np.abs(J).max())
        assert np.abs(Ht).max() == 0.0


@pytest.mark.parametrize("name", NAMES)
def test_start_solutions_solve_start_system(name):
    if name in BAD_START_DATA:
        pytest.skip("the reference's start solutions do not solve this system (reference data)")
    d = load(name)
    op = oracle_problem(d)
    q = d["start_params"]
    hx, ht = oracle.p2c_poly(*d["cmap"], q, q, d["M"])
    worst = 0.0
    for x in d["start_sols"]:
        H, _, _ = op.eval(hx[0], ht[0], x, 0.0)
        worst = max(worst, np.abs(H).max() / max(1.0, np.abs(x).max()))
    # the start solutions are stored with ~16 significant digits
    assert worst < 1e-9, worst


@pytest.mark.parametrize("name", ["P3P", "3view_triangulation", "dual_reciever_TDOA_5pt"])
def test_dhdt_along_the_path(name):
    """dH/dt of the tables equals the central difference of H in t."""
    d = load(name)
    op = oracle_problem(d)
    rng = np.random.default_rng(3)
    p1 = (rng.standard_normal(d["k"]) + 1j * rng.standard_normal(d["k"])) / np.sqrt(2)
    hx, ht = oracle.p2c_poly(*d["cmap"], d["start_params"], p1, d["M"])
    x = d["start_sols"][0]
    t, h = 0.4, 1e-5
    _, _, Ht = op.eval(hx[0], ht[0], x, t)
    Hp, _, _ = op.eval(hx[0], ht[0], x, t + h)
    Hm, _, _ = op.eval(hx[0], ht[0], x, t - h)
    fd = (Hp - Hm) / (2 * h)
    assert np.abs(Ht - fd).max() <= 1e-7 * max(1.0, np.abs(Ht).max())


def test_oracle_tracks_p3p_to_its_target():
    """P3P from its start to its target parameters.  The system has 8 = 2^3
    roots (Bezout), in +-x pairs (every monomial has even degree).  With the
    reference's settings (15 corrector iterations, dt0 = 0.05) two paths jump
    onto one root; with 3 corrector iterations and dt0 = 0.01 the 8 start
    solutions map 1-to-1 onto the 8 roots, which come in +-x pairs."""
    d = load("P3P")
    op = oracle_problem(d)
    hx, ht = oracle.p2c_poly(*d["cmap"], d["start_params"], d["target_params"], d["M"])
    c1, _ = oracle.p2c_poly(*d["cmap"], d["target_params"], d["target_params"], d["M"])
    for kw, distinct in ((dict(max_steps=d["max_steps"], max_corr=d["max_corr"], incr_steps=d["incr_steps"]), False),
                         (dict(max_steps=400, max_corr=3, dt0=0.01), True)):
        res = op.track(d["start_sols"], hx, ht, **kw)
        assert ((res["flags"] & oracle.FLAG_REFINED) != 0).all()
        for x in res["roots"]:
            H, _, _ = op.eval(c1[0], np.zeros_like(ht[0]), x, 1.0)
            assert np.abs(H).max() < 1e-10 * max(1.0, np.abs(x).max()) ** 2
        r = res["roots"]
        dist = np.abs(r[:, None, :] - r[None, :, :]).max(axis=2) + np.eye(8)
        assert (dist.min() > 1e-6) == distinct
        if distinct:
            neg = np.abs(r[:, None, :] + r[None, :, :]).max(axis=2)
            assert (neg.min(axis=1) < 1e-9).all()


def test_reference_converged_roots_fixture():
    d = load("uncalibrated_trifocal_rel_pos_CY")
    assert d["ref_converged"].shape == (444, 12)
    assert d["start_sols"].shape == (668, 12)


SYS_TEXT = """function [f, numOfVars, num_of_params] = sys_demo()
    syms x1 x2
    syms p1 p2 p3
    f(1) = 2*p1*x1^2 + 2*p2*x1*x2 - 3*x2 + p3^2 - 1;
    f(2) = x1*x2*p1 - 4*p2*p3*x2^2 + 6*x1 + 5;
end
"""


def test_compile_system_own_numbering():
    import sympy as sp
    sd = cc.parse_system(SYS_TEXT, "demo")
    comp = cc.compile_system(sd)
    assert comp["n"] == 2 and comp["k"] == 3 and comp["max_order_t"] == 2
    scal, prm = cc.coefficient_map_arrays(comp["coeffs"], sd.ps)
    # the coefficient numbering a compile produces is recovered by inference
    coef = cc.infer_coefficient_map(sd, comp["ht"], comp["hx"], comp["C"])
    assert all(sp.expand(a - b) == 0 for a, b in zip(coef, comp["coeffs"]))
    op = oracle.OracleProblem(2, comp["C"], comp["hx"], comp["ht"], comp["max_order_t"])
    rng = np.random.default_rng(5)
    f = sp.lambdify(sd.xs + sd.ps, sd.fs, "numpy")
    for _ in range(3):
        x = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        p = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        hx, ht = oracle.p2c_poly(scal, prm, p, p, comp["max_order_t"])
        H, _, _ = op.eval(hx[0], ht[0], x, 0.0)
        assert np.allclose(H, np.array(f(*x, *p), dtype=complex), rtol=1e-13, atol=1e-13)


def test_infer_rejects_wrong_tables():
    sd = cc.parse_system(SYS_TEXT, "demo")
    comp = cc.compile_system(sd)
    bad = comp["ht"].copy()
    bad[0, 0, 0] += 1  # wrong scalar on one term
    with pytest.raises(ValueError):
        cc.infer_coefficient_map(sd, bad, comp["hx"], comp["C"])
    badx = comp["hx"].copy()
    i, j, u = np.argwhere(badx[..., 0] != 0)[0]
    badx[i, j, u, 0] *= -1  # wrong sign in dH/dx
    with pytest.raises(ValueError):
        cc.infer_coefficient_map(sd, comp["ht"], badx, comp["C"])


def test_ph_conversion_pairs():
    d = load("trifocal_2op1p_30x30")
    scal, prm = d["cmap"]
    assert d["C"] == scal.shape[0] and d["M"] == 2
    # every coefficient is one product of two parameters (or of one and the constant)
    assert prm.shape[1:] == (1, 2)
    pairs = {tuple(sorted(r[0])) for r in prm}
    assert len(pairs) == d["C"] and (d["k"], d["k"]) not in pairs


def test_reference_text_readers(tmp_path):
    """The C-ABI readers and the compiler's readers agree on files written in
    the reference's tab-separated formats."""
    import hcx
    d = load("P3P")
    n = d["n"]
    (tmp_path / "start_params.txt").write_text(
        "\n".join(f"{float(z.real)!r}\t{float(z.imag)!r}" for z in d["start_params"]))
    (tmp_path / "start_sols.txt").write_text(
        "\n".join(f"{float(z.real)!r}\t{float(z.imag)!r}" for z in d["start_sols"].reshape(-1)) + "\n")
    (tmp_path / "dHdt_indx.txt").write_text(
        "\n".join("\t".join(str(v) for v in row.reshape(-1)) + "\t" for row in d["ht3"]))
    a = cc.read_complex(str(tmp_path / "start_params.txt"))
    b = hcx.ref_read_params(str(tmp_path / "start_params.txt"), cap=64)
    assert np.array_equal(a, d["start_params"]) and np.array_equal(b, a)
    s = hcx.ref_read_start_sols(str(tmp_path / "start_sols.txt"), n, cap_tracks=64)
    assert np.array_equal(s, d["start_sols"])
    t = hcx.ref_read_indices(str(tmp_path / "dHdt_indx.txt"), cap=4096)
    assert np.array_equal(t, d["ht3"].reshape(-1))
