"""Synthetic workloads of the tracker benchmarks (five configurations).

Inputs are reproducible from seeds:
  CN(0,1) = (a + i b)/sqrt(2), a, b ~ N(0,1) (SplitMix64 + Box-Muller in libhcx)
  gamma   = exp(2 pi i u), u ~ U[0,1)
  c_j(p)  = A[j,0] + sum_q A[j,1+q] p_q with A, p ~ CN(0,1)
Every equation's support contains {1, x_1^d, ..., x_n^d}, so a generic target
system has exactly prod d_i roots (Bezout) and the gamma-trick total-degree
homotopy maps start roots 1-to-1 onto target roots.

Start roots of a parameter homotopy come from a seed instance p0 solved by
total-degree homotopy on a seeded subset of the prod d_i start roots; the first
K refined, distinct roots (in subset order) are kept; further rounds (new
subset, new gamma) run until K distinct roots are found.
"""
import numpy as np

from . import (FLAG_REFINED, PREC_FP64, Batch, Problem, System, collect, default_opts, random_cn, random_gamma,
               track)

# (n, paths, extra monomials) of the quadratic problem types used by the mixed
# batch and the serving loop.
TYPE_PATHS = {3: 8, 5: 32, 6: 64, 9: 142, 12: 583, 16: 40, 30: 312}
EXTRA_MONOMIALS = 5


def _seed(base, tag):
    return (int(base) * 1000003 + int(tag) * 7919) & 0xFFFFFFFFFFFFFFFF


def quadratic_type(n, num_params, seed, monomials=None):
    """Sparse quadratic parametric system of one problem type."""
    if monomials is None:
        monomials = n + 1 + EXTRA_MONOMIALS
    return System.synthetic(n, num_params, num_quadratic=n, monomials=monomials, seed=seed)


def total_degree_problem(system, target_coeffs, gamma, root_index):
    """Device problem of (1-t) gamma G + t F over the given start-root indices."""
    td, hx, ht = system.total_degree(target_coeffs, gamma)
    start = td.start_roots(root_index)
    inf = td.info
    hxi, hti = td.tables()
    prob = Problem(inf["num_vars"], inf["num_coeffs"], hxi, hti, start)
    return td, prob, hx, ht, start


def _distinct_mask(roots, ok, tol):
    keep = np.zeros(len(roots), dtype=bool)
    chosen = []
    for j in range(len(roots)):
        if not ok[j]:
            continue
        r = roots[j]
        dup = False
        for c in chosen:
            if np.all(np.abs(r.real - c.real) < tol) and np.all(np.abs(r.imag - c.imag) < tol):
                dup = True
                break
        if not dup:
            keep[j] = True
            chosen.append(r)
    return keep


def seed_start_roots(system, p0, K, seed, opts=None, margin=1.3, max_rounds=4):
    """K start roots of the parameter homotopy at p0 (GPU total-degree solve)."""
    import torch
    if opts is None:
        opts = default_opts(precision=PREC_FP64, max_steps=400)
    coeffs = system.eval_coeffs(p0)
    count = K
    found = np.zeros((0, system.info["num_vars"]), dtype=np.complex128)
    for rnd in range(max_rounds):
        # a new gamma each round: when K is the whole Bezout number a path
        # that fails for one gamma is tracked again along another path
        gamma = random_gamma(_seed(seed, 11) if rnd == 0 else _seed(seed, 1000 + rnd))
        want = int(np.ceil(count * margin)) + 16
        idx = system.subset(want, _seed(seed, 13 + rnd))
        idx = idx[idx >= 0]
        td, prob, hx, ht, start = total_degree_problem(system, coeffs, gamma, idx)
        b = Batch(prob, 1, coeffs_hx=hx.reshape(1, -1, 2), coeffs_ht=ht.reshape(1, -1, 1))
        res = track(prob, b, opts)
        torch.cuda.synchronize()
        h = res.to_host()
        ok = (h["flags"] & FLAG_REFINED) != 0
        allr = np.concatenate([found, h["roots"][ok]])
        keep = _distinct_mask(allr, np.ones(len(allr), dtype=bool), 1e-6)
        found = allr[keep]
        prob.close()
        if len(found) >= K:
            return found[:K]
        count = K * 2
    raise RuntimeError(f"seed_start_roots: found {len(found)} of {K} roots")


def config1(seed=1):
    """Dense quadratic n=3 total-degree homotopy, one instance, 8 paths."""
    sysf = System.synthetic(3, 0, num_quadratic=3, monomials=0, seed=_seed(seed, 1))
    coeffs = sysf.eval_coeffs(None)
    gamma = random_gamma(_seed(seed, 2))
    td, hx, ht = sysf.total_degree(coeffs, gamma)
    start = td.start_roots(np.arange(8))
    return dict(system=sysf, homotopy=td, coeffs_hx=hx.reshape(1, -1, 2), coeffs_ht=ht.reshape(1, -1, 1),
                start=start, gamma=gamma, seed=seed)


def config2_system(seed=2):
    """n=6 parametric system: 12 monomials per equation, k=24 parameters."""
    sysf = System.synthetic(6, 24, num_quadratic=6, monomials=12, seed=_seed(seed, 1))
    p0 = random_cn(_seed(seed, 3), 24)
    return sysf, p0


def targets(k, count, seed):
    return random_cn(seed, count * k).reshape(count, k)


def config4_system(seed=4):
    """n=12: 6 quadratic + 6 cubic equations, 20 monomials each, k=36."""
    sysf = System.synthetic(12, 36, num_quadratic=6, monomials=20, seed=_seed(seed, 1))
    p0 = random_cn(_seed(seed, 3), 36)
    return sysf, p0


CONFIG3_TYPES = [(3, 8), (5, 32), (6, 64), (9, 142), (12, 583)]
CONFIG5_TYPES = [(3, 8), (6, 64), (9, 142), (12, 583), (16, 40), (30, 312)]


def quadratic_types(types, seed, num_params=None):
    """Problem types of the uneven batch / serving workloads: per type a sparse
    quadratic parametric system (k = 2n parameters unless given), its seed
    parameters p0 and K start roots from the seed total-degree solve."""
    out = []
    for ti, (n, K) in enumerate(types):
        k = num_params if num_params else 2 * n
        sysf = quadratic_type(n, k, _seed(seed, 100 + ti))
        p0 = random_cn(_seed(seed, 200 + ti), k)
        start = seed_start_roots(sysf, p0, K, seed=_seed(seed, 300 + ti))
        out.append(dict(n=n, K=K, k=k, system=sysf, p0=p0, start=start))
    return out


def mixed_instances(num_types, count, seed):
    """Type of each instance, uniform at random (SplitMix64 stream)."""
    u = random_cn(_seed(seed, 400), count)
    return (np.floor((np.angle(u) + np.pi) / (2 * np.pi) * num_types).astype(np.int64) % num_types)
