"""GPU tracker (libhcx HIP kernels, called through the C-ABI) against the FP64
oracle on the same systems, parameters, start roots and gamma.

Bar: identical refined (root-level converged) and real counts, identical path
flags, and every refined root within 1e-10 relative error of its oracle root
(FP64 refinement at t = 1 in both)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT_TOL = 1e-10


def rel_err(a, b):
    return np.abs(a - b).max(axis=-1) / np.maximum(1.0, np.abs(b).max(axis=-1))


def compare(hcx, oracle, g, o, label, exact_flags=True):
    mask = hcx.FLAG_TRACKED | hcx.FLAG_INFINITY | hcx.FLAG_REFINED | hcx.FLAG_REAL
    fg, fo = g["flags"] & mask, o["flags"] & mask
    ref_g = (fg & hcx.FLAG_REFINED) != 0
    ref_o = (fo & hcx.FLAG_REFINED) != 0
    print(f"[{label}] refined gpu={ref_g.sum()} oracle={ref_o.sum()} real gpu={((fg & hcx.FLAG_REAL) != 0).sum()} "
          f"oracle={((fo & hcx.FLAG_REAL) != 0).sum()}", flush=True)
    assert ref_g.sum() == ref_o.sum(), label
    assert ((fg & hcx.FLAG_REAL) != 0).sum() == ((fo & hcx.FLAG_REAL) != 0).sum(), label
    if exact_flags:
        assert (fg == fo).all(), f"{label}: flags differ at {np.nonzero(fg != fo)[0][:10]}"
    both = ref_g & ref_o
    err = rel_err(g["roots"][both], o["roots"][both])
    print(f"[{label}] max root rel err {err.max() if len(err) else 0:.3e}", flush=True)
    assert (err <= ROOT_TOL).all(), f"{label}: max rel err {err.max()}"
    return err


def config1_case(hcx, oracle, seed):
    from hcx import workloads
    c = workloads.config1(seed)
    td = c["homotopy"]
    hxi, hti = td.tables()
    inf = td.info
    prob = hcx.Problem(3, inf["num_coeffs"], hxi, hti, c["start"])
    op = oracle.OracleProblem(3, inf["num_coeffs"], hxi, hti, 1)
    return c, prob, op


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("prec", [0, 1])
def test_config1(hcx_mod, oracle_mod, gpu, seed, prec):
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    c, prob, op = config1_case(hcx, oracle, seed)
    opts = hcx.default_opts(precision=prec)
    b = hcx.Batch(prob, 1, coeffs_hx=c["coeffs_hx"], coeffs_ht=c["coeffs_ht"])
    r = hcx.track(prob, b, opts)
    torch.cuda.synchronize()
    g = r.to_host()
    o = op.track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1)
    compare(hcx, oracle, g, o, f"config1 seed={seed} prec={prec}", exact_flags=True)
    assert ((g["flags"] & hcx.FLAG_REFINED) != 0).sum() == 8
    if prec == 0:
        # FP64 end to end: the same step sequence as the oracle
        assert (g["steps"] == o["steps"]).all() and (g["newton"] == o["newton"]).all()


@pytest.fixture(scope="module")
def config2_small(hcx_mod, oracle_mod, gpu):
    from hcx import workloads
    sysf, p0 = workloads.config2_system(2)
    start = workloads.seed_start_roots(sysf, p0, 64, seed=2)
    p1 = workloads.targets(24, 24, seed=77)
    return sysf, p0, start, p1


@pytest.mark.parametrize("prec", [0, 1])
def test_config2_parameter_homotopy(hcx_mod, oracle_mod, config2_small, prec):
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    sysf, p0, start, p1 = config2_small
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    b = hcx.Batch(prob, len(p1), target_params=p1)
    opts = hcx.default_opts(precision=prec)
    r = hcx.track(prob, b, opts)
    torch.cuda.synchronize()
    g = r.to_host()
    hxi, hti = sysf.tables()
    op = oracle.OracleProblem(6, sysf.info["num_coeffs"], hxi, hti, 1)
    ohx, oht = oracle.p2c_affine(sysf.affine(), p0, p1)
    o = op.track(start, ohx, oht, threads=8)
    compare(hcx, oracle, g, o, f"config2 prec={prec}", exact_flags=(prec == 0))
    # most paths of a generic instance reach a root
    assert ((g["flags"] & hcx.FLAG_REFINED) != 0).mean() > 0.9
    gc = hcx.collect(r, opts)
    hc, _ = hcx.collect_host(6, len(p1), 64, g["roots"], g["flags"])
    oc, _ = oracle.collect(6, len(p1), 64, o["roots"], o["flags"])
    assert gc == hc
    assert gc["refined"] == oc["refined"] and gc["real"] == oc["real"] and gc["unique"] == oc["unique"]


@pytest.mark.parametrize("prec", [0, 1])
def test_static_launch_equals_persistent(hcx_mod, config2_small, prec):
    """The reference's launch shape (one path per group, no refill) gives the
    persistent launch's results; in mixed precision both are followed by the
    FP64 pass (escalations, duplicate roots)."""
    import torch
    hcx = hcx_mod
    sysf, p0, start, p1 = config2_small
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    opts = hcx.default_opts(precision=prec)
    b = hcx.Batch(prob, 4, target_params=p1[:4])
    a = hcx.track(prob, b, opts).to_host()
    s = hcx.track(prob, b, opts, static=True).to_host()
    torch.cuda.synchronize()
    assert (a["flags"] == s["flags"]).all()
    assert (a["roots"] == s["roots"]).all()
    assert (a["steps"] == s["steps"]).all()


def test_device_p2c_equals_host_tables(hcx_mod, config2_small):
    import torch
    hcx = hcx_mod
    sysf, p0, start, p1 = config2_small
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    opts = hcx.default_opts()
    I = 6
    hx = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[0] for i in range(I)])
    ht = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[1] for i in range(I)])
    a = hcx.track(prob, hcx.Batch(prob, I, target_params=p1[:I]), opts).to_host()
    b = hcx.track(prob, hcx.Batch(prob, I, coeffs_hx=hx, coeffs_ht=ht), opts).to_host()
    torch.cuda.synchronize()
    assert (a["flags"] == b["flags"]).all()
    assert np.abs(a["roots"] - b["roots"]).max() < 1e-12


def test_lanes_per_path_variants(hcx_mod, config2_small):
    """Power-of-two groups (G=8) and exact groups (G=n=6) agree."""
    import torch
    hcx = hcx_mod
    sysf, p0, start, p1 = config2_small
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    b = hcx.Batch(prob, 4, target_params=p1[:4])
    a = hcx.track(prob, b, hcx.default_opts(lanes_per_path=6)).to_host()
    c = hcx.track(prob, b, hcx.default_opts(lanes_per_path=8)).to_host()
    torch.cuda.synchronize()
    ok = ((a["flags"] & hcx.FLAG_REFINED) != 0) & ((c["flags"] & hcx.FLAG_REFINED) != 0)
    assert ok.sum() >= 0.9 * len(ok)
    assert rel_err(a["roots"][ok], c["roots"][ok]).max() < 1e-10


def test_multi_segment_launch(hcx_mod, oracle_mod, gpu):
    """Three problem types in one persistent launch equal separate launches."""
    import torch
    from hcx import workloads
    hcx = hcx_mod
    probs, batches, singles = [], [], []
    for n, K, seed in [(3, 8, 31), (5, 32, 32), (6, 64, 33)]:
        sysf = workloads.quadratic_type(n, 4, seed)
        p0 = hcx.random_cn(seed + 100, 4)
        start = workloads.seed_start_roots(sysf, p0, K, seed=seed)
        prob = hcx.Problem.from_system(sysf, start, p0=p0)
        p1 = workloads.targets(4, 5, seed + 200)
        probs.append(prob)
        batches.append(hcx.Batch(prob, 5, target_params=p1))
    opts = hcx.default_opts()
    multi = hcx.track_multi(probs, batches, opts)
    torch.cuda.synchronize()
    for prob, b, m in zip(probs, batches, multi):
        s = hcx.track(prob, b, opts).to_host()
        mh = m.to_host()
        torch.cuda.synchronize()
        assert (mh["flags"] == s["flags"]).all()
        assert np.abs(mh["roots"] - s["roots"]).max() < 1e-12
        assert (mh["inst_done"][: b.I] > 0).all()


@pytest.mark.parametrize("prec", [0, 1])
def test_multi_segment_same_size(hcx_mod, gpu, prec):
    """Three different n = 6 problems in one launch run on the size-specialised
    instance, each wave copying the next segment's compact tables into LDS when
    it moves on; every segment equals its separate launch bit for bit."""
    import torch
    from hcx import workloads
    hcx = hcx_mod
    probs, batches = [], []
    for seed in (41, 42, 43):
        sysf = workloads.quadratic_type(6, 12, seed)
        p0 = hcx.random_cn(seed + 100, 12)
        start = workloads.seed_start_roots(sysf, p0, 64, seed=seed)
        prob = hcx.Problem.from_system(sysf, start, p0=p0)
        probs.append(prob)
        batches.append(hcx.Batch(prob, 6, target_params=workloads.targets(12, 6, seed + 200)))
    opts = hcx.default_opts(precision=prec)
    multi = hcx.track_multi(probs, batches, opts)
    torch.cuda.synchronize()
    # replicas = 1: the separate launch keeps whole term sums per lane, as
    # the multi-segment launch does
    one_opts = hcx.default_opts(precision=prec, replicas=1)
    for prob, b, m in zip(probs, batches, multi):
        s = hcx.track(prob, b, one_opts).to_host()
        mh = m.to_host()
        torch.cuda.synchronize()
        assert (mh["flags"] == s["flags"]).all()
        assert (mh["steps"] == s["steps"]).all()
        assert (mh["roots"] == s["roots"]).all()


def test_reference_shaped_launcher(hcx_mod, oracle_mod, gpu):
    """hcx_kernel_HC_Solver: [K][n+1] buffers, bool flags, (t, dt) debug, elapsed seconds."""
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    c, prob, op = config1_case(hcx, oracle, 3)
    K, n = 8, 3
    start = np.concatenate([c["start"], np.ones((K, 1))], axis=1)
    d_start = torch.from_numpy(start).cuda()
    d_track = d_start.clone()
    hx = torch.from_numpy(c["coeffs_hx"].reshape(-1)).cuda()
    ht = torch.from_numpy(c["coeffs_ht"].reshape(-1)).cuda()
    conv = torch.zeros(K, dtype=torch.uint8, device="cuda")
    inf = torch.zeros(K, dtype=torch.uint8, device="cuda")
    dbg = torch.zeros(K, dtype=torch.complex128, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    secs = hcx.lib().hcx_kernel_HC_Solver(stream, prob.handle, 100, 3, 3, C.c_void_p(d_start.data_ptr()),
                                          C.c_void_p(d_track.data_ptr()), C.c_void_p(hx.data_ptr()),
                                          C.c_void_p(ht.data_ptr()), C.c_void_p(conv.data_ptr()),
                                          C.c_void_p(inf.data_ptr()), C.c_void_p(dbg.data_ptr()))
    assert secs > 0, hcx.lib().hcx_last_error()
    o = op.track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1, max_steps=100, dt_min=0.0,
                 refine_iters=0)
    tr = d_track.cpu().numpy()
    assert (tr[:, n] == 1).all()
    assert (conv.cpu().numpy() == ((o["flags"] & oracle.FLAG_TRACKED) != 0)).all()
    assert np.abs(tr[:, :n] - o["roots"]).max() < 1e-12
    assert (dbg.cpu().numpy()[conv.cpu().numpy() == 1] == 1).all()
