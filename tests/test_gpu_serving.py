"""Serving loop: continuous batching over resident problem types.  A request's
result is the same alone and in a mixed stream; the HBM budget gates
admission; results match the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(hcx_mod, gpu):
    from hcx import workloads
    hcx = hcx_mod
    types = workloads.quadratic_types([(3, 8), (6, 64), (9, 142)], seed=9)
    probs = [hcx.Problem.from_system(t["system"], t["start"], p0=t["p0"]) for t in types]
    return types, probs


def wait(server, rid, tid, timeout=60.0):
    import time
    t0 = time.time()
    while True:
        r = server.result(rid, tid)
        if r is not None:
            return r
        assert time.time() - t0 < timeout, "request did not finish"
        time.sleep(0.001)


@pytest.mark.parametrize("prec", [1, 0])
def test_alone_equals_in_mix(hcx_mod, served, prec):
    """A request's result is the same bit for bit alone and in a mix, in
    mixed precision and in FP64 (where a lone request's launch holds few
    paths and runs the latency instance, a mixed batch the throughput one)."""
    from hcx import workloads
    hcx = hcx_mod
    types, probs = served
    srv = hcx.Server(hcx.default_opts(precision=prec), hbm_budget_bytes=256 << 20)
    tids = [srv.add_type(p) for p in probs]
    x = workloads.targets(types[2]["k"], 1, 123)[0]
    alone = wait(srv, srv.submit(tids[2], x), tids[2])
    rng = np.random.default_rng(0)
    ids = []
    for i in range(60):
        t = int(rng.integers(0, 3))
        ids.append((srv.submit(tids[t], workloads.targets(types[t]["k"], 1, 1000 + i)[0]), t))
        if i == 30:
            mid = srv.submit(tids[2], x)
    srv.drain()
    mix = wait(srv, mid, tids[2])
    for rid, t in ids:
        assert wait(srv, rid, t) is not None
    st = srv.stats()
    print("serving stats", st, flush=True)
    assert (alone[2] == mix[2]).all()
    assert np.array_equal(alone[1], mix[1]), "result changed between alone and in the mix"
    assert st["completed"] == st["submitted"] == 62 and st["launches"] >= 2
    srv.close()


def test_budget_gates_admission(hcx_mod, served):
    from hcx import workloads
    hcx = hcx_mod
    types, probs = served
    # room for only a few n=9 requests at a time
    srv = hcx.Server(hcx.default_opts(precision=hcx.PREC_MIXED), hbm_budget_bytes=1 << 20, max_batch_requests=8)
    tid = srv.add_type(probs[2])
    ids = [srv.submit(tid, workloads.targets(types[2]["k"], 1, 50 + i)[0]) for i in range(24)]
    srv.drain()
    st = srv.stats()
    assert st["completed"] == 24 and st["launches"] >= 3
    assert st["peak_inflight_paths"] <= 24 * 142
    for rid in ids:
        assert wait(srv, rid, tid) is not None
    srv.close()


def test_served_roots_match_oracle(hcx_mod, oracle_mod, served):
    from hcx import workloads
    hcx, oracle = hcx_mod, oracle_mod
    types, probs = served
    srv = hcx.Server(hcx.default_opts(precision=hcx.PREC_MIXED))
    tids = [srv.add_type(p) for p in probs]
    for ti, ty in enumerate(types):
        x = workloads.targets(ty["k"], 1, 777 + ti)
        info, roots, flags = wait(srv, srv.submit(tids[ti], x[0]), tids[ti])
        sysf = ty["system"]
        hxi, hti = sysf.tables()
        op = oracle.OracleProblem(ty["n"], sysf.info["num_coeffs"], hxi, hti, 1)
        ohx, oht = oracle.p2c_affine(sysf.affine(), ty["p0"], x)
        o = op.track(ty["start"], ohx, oht, threads=8)
        rg, ro = (flags & 4) != 0, (o["flags"] & 4) != 0
        assert rg.sum() == ro.sum() == info["refined"]
        ok = rg & ro
        err = np.abs(roots[ok] - o["roots"][ok]).max(axis=1) / np.maximum(1, np.abs(o["roots"][ok]).max(axis=1))
        assert err.max() <= 1e-10
    srv.close()
