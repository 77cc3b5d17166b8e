"""Kernel instances that must agree bit for bit.

The size-specialised instances read compact term tables from LDS (n <= 16)
and reduce pivot candidates with DPP permutations (power-of-two lanes per
path, n = 30 on 32-lane groups); the generic instances read the full tables
from global memory and reduce with lane exchanges.  Both evaluate the same
terms in the same order and pick the same pivots, so with the same lanes per
path every output (flags, step and Newton counts, roots) is identical.  With
different lanes per path only the group sums of the Newton norms change their
order, so roots agree to the parity tolerance instead."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _type(hcx, n, K, seed=5):
    from hcx import workloads
    ty = workloads.quadratic_types([(n, K)], seed=seed)[0]
    prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
    return ty, prob


def _run(hcx, prob, batch, **kw):
    import torch
    # whole term sums per lane unless a test asks for replica groups (their
    # split sums round differently)
    kw.setdefault("replicas", 1)
    r = hcx.track(prob, batch, hcx.default_opts(**kw))
    torch.cuda.synchronize()
    return r.to_host()


def _assert_identical(a, b, label):
    print(f"[{label}] refined {int(((a['flags'] & 4) != 0).sum())} / {len(a['flags'])}", flush=True)
    assert (a["flags"] == b["flags"]).all(), f"{label}: flags differ at {np.nonzero(a['flags'] != b['flags'])[0][:8]}"
    assert (a["steps"] == b["steps"]).all(), label
    assert (a["newton"] == b["newton"]).all(), label
    both_nan = np.isnan(a["roots"]) & np.isnan(b["roots"])
    assert ((a["roots"] == b["roots"]) | both_nan).all(), f"{label}: roots differ"


@pytest.mark.parametrize("n,K,instances", [(6, 64, 16), (12, 583, 2), (16, 40, 16)])
@pytest.mark.parametrize("prec", [0, 1])
def test_lds_tables_equal_global_tables(hcx_mod, gpu, n, K, instances, prec):
    """Specialised instance (LDS compact tables) == generic instance (global
    tables), same lanes per path."""
    from hcx import workloads
    hcx = hcx_mod
    ty, prob = _type(hcx, n, K)
    b = hcx.Batch(prob, instances, target_params=workloads.targets(ty["k"], instances, seed=77))
    a = _run(hcx, prob, b, precision=prec, lanes_per_path=n, specialize=1)
    g = _run(hcx, prob, b, precision=prec, lanes_per_path=n, specialize=0)
    _assert_identical(a, g, f"n={n} prec={prec}")
    prob.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_n30_dpp_groups(hcx_mod, gpu, prec):
    """n = 30: the specialised 32-lane instance (DPP argmax) equals the generic
    instance at 32 lanes (lane-exchange argmax) exactly, and the generic
    instance at 30 lanes to the parity tolerance (a path whose step sequence
    changes with the summation order may end on another root, so 99% of the
    paths refined by both must agree)."""
    from hcx import workloads
    hcx = hcx_mod
    ty, prob = _type(hcx, 30, 312)
    b = hcx.Batch(prob, 2, target_params=workloads.targets(ty["k"], 2, seed=77))
    a = _run(hcx, prob, b, precision=prec, specialize=1)
    g32 = _run(hcx, prob, b, precision=prec, lanes_per_path=32, specialize=0)
    _assert_identical(a, g32, f"n=30 G=32 prec={prec}")
    g30 = _run(hcx, prob, b, precision=prec, lanes_per_path=30, specialize=0)
    ra, rb = (a["flags"] & 4) != 0, (g30["flags"] & 4) != 0
    assert ra.sum() == rb.sum()
    both = ra & rb
    err = np.abs(a["roots"][both] - g30["roots"][both]).max(axis=1) / np.maximum(
        1.0, np.abs(g30["roots"][both]).max(axis=1))
    agree = int((err <= 1e-10).sum())
    print(f"[n=30 G=30 prec={prec}] refined {int(ra.sum())}, agree {agree} / {int(both.sum())}", flush=True)
    assert both.sum() >= 0.99 * ra.sum()
    assert agree >= 0.99 * both.sum()
    prob.close()


@pytest.mark.parametrize("generic", [False, True])
@pytest.mark.parametrize("replicas", [2, 4, 8])
def test_replica_groups_match_whole_sums(hcx_mod, gpu, oracle_mod, replicas, generic):
    """FP64 launches of few paths split every term sum of the evaluation over
    R replica groups of each path (each replica sums every R-th term, a
    butterfly adds the partial sums).  Only the summation order changes: the
    flags and step counts of every path equal those of the launch with whole
    sums per lane (replicas = 1), and the refined roots agree to 1e-10
    relative, as they do with the oracle's.  Size-specialised instance (6
    lanes per replica) and generic instance (8 lanes per replica, two of
    them without a row)."""
    import torch
    from hcx import workloads
    hcx, oracle = hcx_mod, oracle_mod
    ty = workloads.quadratic_types([(6, 64)], seed=9)[0]
    prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
    b = hcx.Batch(prob, 2, target_params=workloads.targets(ty["k"], 2, seed=78))
    shape = dict(specialize=0, lanes_per_path=8) if generic else {}
    one = _run(hcx, prob, b, precision=0, replicas=1, **shape)
    rep = _run(hcx, prob, b, precision=0, replicas=replicas, **shape)
    assert (one["flags"] == rep["flags"]).all()
    assert (one["steps"] == rep["steps"]).all()
    ok = (one["flags"] & hcx.FLAG_REFINED) != 0
    assert ok.mean() > 0.9
    err = np.abs(one["roots"][ok] - rep["roots"][ok]).max(axis=1) / np.maximum(
        1.0, np.abs(one["roots"][ok]).max(axis=1))
    print(f"[replicas={replicas} generic={generic}] refined {int(ok.sum())}, max rel diff {err.max():.2e}",
          flush=True)
    assert err.max() <= 1e-10
    # and against the oracle on the same paths
    sysf = ty["system"]
    hxi, hti = sysf.tables()
    op = oracle.OracleProblem(6, sysf.info["num_coeffs"], hxi, hti, 1)
    ohx, oht = oracle.p2c_affine(sysf.affine(), ty["p0"], b.target_params.cpu().numpy())
    o = op.track(ty["start"], ohx, oht, threads=8)
    ro = (o["flags"] & hcx.FLAG_REFINED) != 0
    assert (ro == ((rep["flags"] & hcx.FLAG_REFINED) != 0)).all()
    both = ro & ok
    e2 = np.abs(rep["roots"][both] - o["roots"][both]).max(axis=1) / np.maximum(
        1.0, np.abs(o["roots"][both]).max(axis=1))
    assert e2.max() <= 1e-10
    torch.cuda.synchronize()
    prob.close()


@pytest.mark.parametrize("name", ["distorted_2view_triangulation", "optimal_PnP_quat"])
def test_degree4_instance_equals_generic(hcx_mod, gpu, name):
    """Terms of degree 4 at n = 5 and n = 4 (two of the reference's problems):
    the size-specialised degree-4 instance (8-byte compact terms in LDS) and
    the generic instance (degree-6 records whose missing factors read the
    constant 1) give the same flags and step counts and the same roots to
    1e-12 relative, in the throughput and the latency instances."""
    from test_reference_problems import load
    hcx = hcx_mod
    d = load(name)
    prob = hcx.Problem(d["n"], d["C"], d["hx4"], d["ht3"], d["start_sols"], max_order_t=d["M"], cmap=d["cmap"],
                       p0=d["start_params"])
    tp = np.vstack([d["target_params"].reshape(1, -1), hcx.random_cn(5, 2 * d["k"]).reshape(2, -1)])
    b = hcx.Batch(prob, 3, target_params=tp)
    kw = dict(precision=0, max_steps=d["max_steps"], max_corr=3, incr_steps=d["incr_steps"])
    for reps in (1, 0):
        a = _run(hcx, prob, b, specialize=1, replicas=reps, **kw)
        g = _run(hcx, prob, b, specialize=0, lanes_per_path=d["n"], replicas=reps, **kw)
        assert (a["flags"] == g["flags"]).all(), (name, reps)
        assert (a["steps"] == g["steps"]).all(), (name, reps)
        ok = (a["flags"] & hcx.FLAG_REFINED) != 0
        assert ok.mean() > 0.5
        err = np.abs(a["roots"][ok] - g["roots"][ok]).max(axis=1) / np.maximum(
            1.0, np.abs(g["roots"][ok]).max(axis=1))
        print(f"[{name} replicas={reps}] refined {int(ok.sum())}, max rel diff {err.max():.2e}", flush=True)
        assert err.max() <= 1e-12
    prob.close()


def test_replicas_option_checked(hcx_mod, gpu):
    """replicas must be a power of two whose groups fit a wave."""
    from hcx import workloads
    hcx = hcx_mod
    ty, prob = _type(hcx, 6, 64)
    b = hcx.Batch(prob, 1, target_params=workloads.targets(ty["k"], 1, seed=5))
    with pytest.raises(hcx.HcxError, match="invalid tracker options"):
        hcx.track(prob, b, hcx.default_opts(replicas=3))
    with pytest.raises(hcx.HcxError, match="exceed a wave"):
        hcx.track(prob, b, hcx.default_opts(replicas=16))
    prob.close()
