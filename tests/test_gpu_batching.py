"""Uneven multi-type batches: HIP-graph replays and autotuned launch shapes,
checked against eager launches and the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def types(hcx_mod, gpu):
    from hcx import workloads
    return workloads.quadratic_types([(3, 8), (5, 32), (6, 64)], seed=5)


def make_batches(hcx, types, counts, seed):
    from hcx import workloads
    probs, batches = [], []
    for ty, cnt in zip(types, counts):
        prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
        p1 = workloads.targets(ty["k"], cnt, seed + ty["n"])
        probs.append(prob)
        batches.append(hcx.Batch(prob, cnt, target_params=p1))
    return probs, batches


def test_graph_replays_match_eager(hcx_mod, types):
    import torch
    hcx = hcx_mod
    probs, batches = make_batches(hcx, types, [7, 5, 3], 11)
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    eager = [r.to_host() for r in hcx.track_multi(probs, batches, opts)]
    torch.cuda.synchronize()
    g = hcx.Graph(probs, batches, opts)
    for rep in range(3):
        for r in g.results:  # poison outputs so a replay that does nothing fails
            r.roots.fill_(float("nan"))
            r.flags.fill_(0)
        g.launch()
        torch.cuda.synchronize()
        for e, r in zip(eager, g.results):
            h = r.to_host()
            assert (h["flags"] == e["flags"]).all(), f"replay {rep}"
            assert (h["roots"] == e["roots"]).all(), f"replay {rep}"
        # eager work between replays
        hcx.track_multi(probs, batches, opts)
        torch.cuda.synchronize()
    g.close()


def test_autotune_keeps_results(hcx_mod, oracle_mod, types):
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    ty = types[2]  # n = 6
    probs, batches = make_batches(hcx, [ty], [16], 21)
    prob, b = probs[0], batches[0]
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    tr = hcx.autotune(prob, b, opts)
    print("autotune", tr, flush=True)
    assert tr["candidates"] >= 2 and tr["lanes_per_path"] >= 6 and tr["best_ms"] > 0
    g = hcx.track(prob, b, opts).to_host()
    torch.cuda.synchronize()
    sysf = ty["system"]
    hxi, hti = sysf.tables()
    op = oracle.OracleProblem(6, sysf.info["num_coeffs"], hxi, hti, 1)
    p1 = b.target_params.cpu().numpy()
    ohx, oht = oracle.p2c_affine(sysf.affine(), ty["p0"], p1)
    o = op.track(ty["start"], ohx, oht, threads=8)
    ok = ((g["flags"] & 4) != 0) & ((o["flags"] & 4) != 0)
    assert ((g["flags"] & 4) != 0).sum() == ((o["flags"] & 4) != 0).sum()
    err = np.abs(g["roots"][ok] - o["roots"][ok]).max(axis=1) / np.maximum(1, np.abs(o["roots"][ok]).max(axis=1))
    assert err.max() <= 1e-10
    hcx.lib().hcx_problem_set_launch_shape(prob.handle, 0, 1, 0)


def test_uneven_multi_type_vs_oracle(hcx_mod, oracle_mod, types):
    """One launch over three types with uneven instance counts equals the oracle per type."""
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    probs, batches = make_batches(hcx, types, [9, 4, 2], 31)
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    res = hcx.track_multi(probs, batches, opts)
    torch.cuda.synchronize()
    for ty, b, r in zip(types, batches, res):
        g = r.to_host()
        sysf = ty["system"]
        hxi, hti = sysf.tables()
        op = oracle.OracleProblem(ty["n"], sysf.info["num_coeffs"], hxi, hti, 1)
        ohx, oht = oracle.p2c_affine(sysf.affine(), ty["p0"], b.target_params.cpu().numpy())
        o = op.track(ty["start"], ohx, oht, threads=8)
        rg, ro = (g["flags"] & 4) != 0, (o["flags"] & 4) != 0
        assert rg.sum() == ro.sum(), ty["n"]
        ok = rg & ro
        err = np.abs(g["roots"][ok] - o["roots"][ok]).max(axis=1) / np.maximum(1, np.abs(o["roots"][ok]).max(axis=1))
        assert err.max() <= 1e-10, ty["n"]


@pytest.mark.parametrize("prec", [0, 1])
def test_multi_size_launch_equals_single_type_launches(hcx_mod, types, prec):
    """A launch over types of different sizes runs each segment's
    size-specialised loop (the multi-size instance): flags, step and Newton
    counts and roots equal those of one launch per type, bit for bit; the
    generic instance (specialize = 0) gives the same counts."""
    import torch
    hcx = hcx_mod
    probs, batches = make_batches(hcx, types, [9, 6, 4], 31)
    opts = hcx.default_opts(precision=prec)
    multi = [r.to_host() for r in hcx.track_multi(probs, batches, opts)]
    torch.cuda.synchronize()
    # one segment of few paths may split its term sums over replica groups
    # (different rounding); replicas = 1 keeps the sums of the multi-size launch
    one_opts = hcx.default_opts(precision=prec, replicas=1)
    for p, b, m in zip(probs, batches, multi):
        one = hcx.track(p, b, one_opts).to_host()
        torch.cuda.synchronize()
        for k in ("flags", "steps", "newton", "refine"):
            assert (one[k] == m[k]).all(), (p.n, k)
        both_nan = np.isnan(one["roots"]) & np.isnan(m["roots"])
        assert ((one["roots"] == m["roots"]) | both_nan).all(), p.n
    gen = [r.to_host() for r in hcx.track_multi(probs, batches, hcx.default_opts(precision=prec, specialize=0))]
    torch.cuda.synchronize()
    for g, m in zip(gen, multi):
        assert (g["flags"] == m["flags"]).all() and (g["steps"] == m["steps"]).all()
