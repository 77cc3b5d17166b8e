"""Monodromy start-root generation on the GPU tracker.

Replaces the reference's offline start-solution generator, which calls
HomotopyContinuation.jl's `monodromy_solve` (start-sys-gen/P3P_monodromySolve.jl:14-16)
and writes the start parameters and solutions a problem folder ships.  The
method (published: Duff, Hill, Jensen, Lee, Leykin, Sommars, "Solving
polynomial systems via homotopy continuation and monodromy", IMA J. Numer.
Anal. 2019):

1. Start pair.  For coefficients affine in the parameters, c(p) = A0 + A1 p,
   the system F(x; p) = sum_j c_j(p) x^e_j is linear in p at a fixed x: pick
   x0 ~ CN(0,1)^n and solve the n x k linear system F(x0; p) = 0 for p0
   (least-norm solution plus a random null-space component, so that p0 is
   generic).  x0 is then a root of F(x; p0).
2. Loops.  Pick random parameter points p1, p2 and track every known root of
   F(x; p0) along the triangle p0 -> p1 -> p2 -> p0 (three parameter
   homotopies p(t) = (1 - t) pa + t pb, each ending with FP64 refinement).
   A root that comes back as a root not yet known is added; new roots are
   sent around the loop in turn.  When a loop gives nothing new for every
   known root, a fresh triangle is drawn.
3. Stop when the known roots reach `target` (for example the Bezout number of
   a system whose supports contain {1, x_i^d}) or after `stall_loops`
   consecutive triangles without a new root.

Every leg is one batched launch of `hcx.track` (one instance, the pending
roots as start solutions), so all pending roots of a loop move together.
"""
import numpy as np

from . import FLAG_REFINED, PREC_FP64, Batch, Problem, default_opts, random_cn, track


def _seed(base, tag):
    return (int(base) * 1000003 + int(tag) * 7919) & 0xFFFFFFFFFFFFFFFF


def _monomial_values(exps, x):
    return np.prod(np.power(x[None, :], exps), axis=1)


def find_start_pair(system, seed):
    """(x0, p0) with F(x0; p0) = 0 for a system with parameter-affine
    coefficients (System.affine(): c = A[:, 0] + A[:, 1:] p)."""
    inf = system.info
    n, k = inf["num_vars"], inf["num_params"]
    if k < n:
        raise ValueError(f"a start pair needs at least n = {n} parameters (the system has {k})")
    exps, eq = system.monomials()
    A = system.affine()
    x0 = random_cn(_seed(seed, 1), n)
    m = _monomial_values(exps, x0)
    M = np.zeros((n, k), dtype=np.complex128)
    rhs = np.zeros(n, dtype=np.complex128)
    for j in range(len(eq)):
        M[eq[j]] += m[j] * A[j, 1:]
        rhs[eq[j]] -= m[j] * A[j, 0]
    p_ls, *_ = np.linalg.lstsq(M, rhs, rcond=None)
    # random point of the solution space: add a null-space component
    vh = np.linalg.svd(M)[2]
    null = vh[n:].conj().T
    z = random_cn(_seed(seed, 2), null.shape[1]) if null.shape[1] else np.zeros(0, dtype=np.complex128)
    p0 = p_ls + null @ z
    resid = np.abs(M @ p0 - rhs).max() / max(1.0, np.abs(rhs).max())
    if not resid < 1e-10:
        raise RuntimeError(f"find_start_pair: linear solve residual {resid:.3e}")
    return x0, p0


def _distinct(roots, known, tol):
    """Indices of `roots` that differ from every known root and from each
    other (componentwise |re|, |im| within tol * max(1, |x|) counts as equal)."""
    out = []
    pool = list(known)
    for j, r in enumerate(roots):
        s = tol * max(1.0, float(np.abs(r).max()))
        if any(np.all(np.abs(r.real - q.real) < s) and np.all(np.abs(r.imag - q.imag) < s) for q in pool):
            continue
        pool.append(r)
        out.append(j)
    return out


def _track_leg(system, roots, pa, pb, opts):
    """Tracks `roots` of F(x; pa) to F(x; pb); returns (refined roots, mask)."""
    import torch
    prob = Problem.from_system(system, roots, p0=pa)
    try:
        b = Batch(prob, 1, target_params=np.asarray(pb, dtype=np.complex128)[None, :])
        r = track(prob, b, opts)
        torch.cuda.synchronize()
        h = r.to_host()
    finally:
        prob.close()
    ok = (h["flags"] & FLAG_REFINED) != 0
    return h["roots"], ok


def monodromy_solve(system, seed, target=None, stall_loops=3, max_loops=50, tol=1e-6, opts=None, log=None):
    """Roots of F(x; p0) at a generic p0 by monodromy.  Returns (p0, roots
    [K, n]); K = target when given and reached, else the count at which
    `stall_loops` consecutive triangles found nothing new."""
    if opts is None:
        opts = default_opts(precision=PREC_FP64, max_steps=2000)
    x0, p0 = find_start_pair(system, seed)
    k = system.info["num_params"]
    roots = [x0]
    pending = [0]
    stalls = 0
    for loop in range(max_loops):
        p1 = random_cn(_seed(seed, 100 + 2 * loop), k)
        p2 = random_cn(_seed(seed, 101 + 2 * loop), k)
        new_in_loop = 0
        # send pending roots around this triangle until it yields nothing new
        while pending:
            cur = np.array([roots[i] for i in pending])
            for pa, pb in ((p0, p1), (p1, p2), (p2, p0)):
                if len(cur) == 0:
                    break
                out, ok = _track_leg(system, cur, pa, pb, opts)
                cur = out[ok]
            fresh = _distinct(cur, roots, tol)
            pending = []
            for j in fresh:
                roots.append(cur[j])
                pending.append(len(roots) - 1)
            new_in_loop += len(fresh)
            if log:
                log(f"monodromy: loop {loop}: {len(roots)} roots (+{len(fresh)})")
            if target is not None and len(roots) >= target:
                return p0, np.array(roots[:target])
        stalls = stalls + 1 if new_in_loop == 0 else 0
        if stalls >= stall_loops:
            break
        # a fresh triangle sends every known root around again
        pending = list(range(len(roots)))
    return p0, np.array(roots)
