"""Monodromy start-root generation (hcx.monodromy) against what the
mathematics fixes: a generic system of the n = 6 parameter homotopy family
(supports contain {1, x_i^2}) has exactly 2^6 = 64 roots, so monodromy must
find 64 distinct roots of F(x; p0), each solving the system, and the set must
equal the total-degree solve of the same F(x; p0)."""
import numpy as np
import pytest


def _residual(system, p, roots):
    exps, eq = system.monomials()
    c = system.eval_coeffs(p)
    n = system.info["num_vars"]
    out = []
    for x in roots:
        m = np.prod(np.power(x[None, :], exps), axis=1)
        F = np.zeros(n, dtype=np.complex128)
        np.add.at(F, eq, c * m)
        out.append(np.abs(F).max() / max(1.0, np.abs(x).max() ** 2))
    return np.array(out)


def test_start_pair_solves_the_system(hcx_mod):
    """CPU: the start pair (x0, p0) is a root / parameter pair and p0 is not
    the least-norm solution (a random null-space component is added)."""
    from hcx import monodromy, workloads
    sysf, _ = workloads.config2_system(2)
    x0, p0 = monodromy.find_start_pair(sysf, seed=7)
    assert _residual(sysf, p0, [x0]).max() < 1e-12
    x1, p1 = monodromy.find_start_pair(sysf, seed=8)
    assert np.abs(p0 - p1).max() > 1e-3


@pytest.mark.gpu
def test_monodromy_finds_all_64_roots(hcx_mod, gpu):
    from hcx import monodromy, workloads
    sysf, _ = workloads.config2_system(2)
    p0, roots = monodromy.monodromy_solve(sysf, seed=7, target=64, log=print)
    print(f"[monodromy] {len(roots)} roots", flush=True)
    assert len(roots) == 64
    res = _residual(sysf, p0, roots)
    print(f"[monodromy] max residual {res.max():.3e}", flush=True)
    assert res.max() < 1e-9
    # the total-degree solve of the same system finds the same 64 roots
    td = workloads.seed_start_roots(sysf, p0, 64, seed=3)
    d = np.abs(roots[:, None, :] - td[None, :, :]).max(axis=2)
    match = d.min(axis=1)
    print(f"[monodromy] max distance to the total-degree roots {match.max():.3e}", flush=True)
    assert match.max() < 1e-8
    assert len(set(d.argmin(axis=1).tolist())) == 64
