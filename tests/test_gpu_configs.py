"""The benchmark workloads' own shapes under oracle parity (libhcx HIP
kernels through the C-ABI against the FP64 oracle):

* n = 12 systems of 6 quadratic + 6 cubic equations, 20 monomials each,
  k = 36, 512 start roots: mixed precision and FP64 against the oracle, and
  mixed against FP64 on 128 instances (65,536 paths);
* the five quadratic types (n, K) = (3, 8), (5, 32), (6, 64), (9, 142),
  (12, 583) in ONE graph-captured persistent launch, each type against the
  oracle, every replay equal to the eager launch;
* a serving loop carrying all six types n in {3, 6, 9, 12, 16, 30}: every
  type against the oracle, and requests of the n = 12, 16 and 30 types equal
  alone and in a mix.

Bar: identical refined and real counts (and the reference's real_ref), every
refined root within 1e-10 relative of its oracle root; FP64 runs also have
identical flags."""
import time

import numpy as np
import pytest

from test_gpu_parity import compare, rel_err

pytestmark = pytest.mark.gpu


def oracle_for(oracle, ty_system, n, p0, p1, start, threads=16, **kw):
    hxi, hti = ty_system.tables()
    op = oracle.OracleProblem(n, ty_system.info["num_coeffs"], hxi, hti, 1)
    ohx, oht = oracle.p2c_affine(ty_system.affine(), p0, p1)
    return op.track(start, ohx, oht, threads=threads, **kw)


def counts_equal(hcx, g, o, label):
    for flag, name in ((hcx.FLAG_REFINED, "refined"), (hcx.FLAG_REAL, "real"), (hcx.FLAG_REAL_REF, "real_ref"),
                       (hcx.FLAG_TRACKED, "tracked")):
        a, b = int(((g["flags"] & flag) != 0).sum()), int(((o["flags"] & flag) != 0).sum())
        assert a == b, f"{label}: {name} gpu={a} oracle={b}"


# ------------------------------------------------------------------ n = 12 quadratic + cubic
@pytest.fixture(scope="module")
def cubic12(hcx_mod, gpu):
    from hcx import workloads
    print("[cubic12] seeding 512 start roots", flush=True)
    sysf, p0 = workloads.config4_system(4)
    start = workloads.seed_start_roots(sysf, p0, 512, seed=4)
    return sysf, p0, start


@pytest.mark.parametrize("prec", [0, 1])
def test_cubic12_vs_oracle(hcx_mod, oracle_mod, cubic12, prec):
    """Mixed precision runs the size-specialised DEG = 3 instance and the
    FP64 re-track pass after the launch; FP64 runs the FP64 instance."""
    import torch
    from hcx import workloads
    hcx, oracle = hcx_mod, oracle_mod
    sysf, p0, start = cubic12
    p1 = workloads.targets(36, 4, seed=4 * 11 + 3)
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    r = hcx.track(prob, hcx.Batch(prob, 4, target_params=p1), hcx.default_opts(precision=prec))
    torch.cuda.synchronize()
    g = r.to_host()
    t0 = time.time()
    o = oracle_for(oracle, sysf, 12, p0, p1, start)
    print(f"[cubic12 prec={prec}] oracle {time.time() - t0:.1f} s", flush=True)
    compare(hcx, oracle, g, o, f"cubic12 prec={prec}", exact_flags=(prec == 0))
    counts_equal(hcx, g, o, f"cubic12 prec={prec}")
    assert ((g["flags"] & hcx.FLAG_REFINED) != 0).mean() > 0.9
    if prec == 0:
        assert (g["steps"] == o["steps"]).all() and (g["newton"] == o["newton"]).all()


def test_cubic12_mixed_equals_fp64_at_scale(hcx_mod, cubic12):
    """128 instances x 512 paths: the mixed tracker's refined and real counts
    and roots are the FP64 tracker's."""
    import torch
    from hcx import workloads
    hcx = hcx_mod
    sysf, p0, start = cubic12
    p1 = workloads.targets(36, 128, seed=901)
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    b = hcx.Batch(prob, 128, target_params=p1)
    gm = hcx.track(prob, b, hcx.default_opts(precision=hcx.PREC_MIXED)).to_host()
    g6 = hcx.track(prob, b, hcx.default_opts(precision=hcx.PREC_FP64)).to_host()
    torch.cuda.synchronize()
    esc = int(((gm["flags"] & hcx.FLAG_ESCALATED) != 0).sum())
    print(f"[cubic12 x128] escalated/re-tracked in FP64: {esc}", flush=True)
    counts_equal(hcx, gm, g6, "cubic12 x128 mixed vs fp64")
    both = ((gm["flags"] & hcx.FLAG_REFINED) != 0) & ((g6["flags"] & hcx.FLAG_REFINED) != 0)
    assert both.sum() == ((g6["flags"] & hcx.FLAG_REFINED) != 0).sum()
    err = rel_err(gm["roots"][both], g6["roots"][both])
    print(f"[cubic12 x128] max root rel err mixed vs fp64 {err.max():.3e}", flush=True)
    assert err.max() <= 1e-10


# ------------------------------------------------------------------ five types, one graph
CONFIG3 = [(3, 8), (5, 32), (6, 64), (9, 142), (12, 583)]


@pytest.fixture(scope="module")
def five_types(hcx_mod, gpu):
    from hcx import workloads
    print("[five types] seeding start roots", flush=True)
    return workloads.quadratic_types(CONFIG3, seed=3)


def test_five_types_one_graph_vs_oracle(hcx_mod, oracle_mod, five_types):
    import torch
    from hcx import workloads
    hcx, oracle = hcx_mod, oracle_mod
    counts = [5, 4, 3, 2, 2]
    probs, batches, p1s = [], [], []
    for ty, cnt in zip(five_types, counts):
        prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
        p1 = workloads.targets(ty["k"], cnt, seed=300 + ty["n"])
        probs.append(prob)
        p1s.append(p1)
        batches.append(hcx.Batch(prob, cnt, target_params=p1))
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    eager = [r.to_host() for r in hcx.track_multi(probs, batches, opts)]
    torch.cuda.synchronize()
    g = hcx.Graph(probs, batches, opts)
    for rep in range(2):
        for r in g.results:
            r.roots.fill_(float("nan"))
            r.flags.fill_(0)
        g.launch()
        torch.cuda.synchronize()
        for e, r in zip(eager, g.results):
            h = r.to_host()
            assert (h["flags"] == e["flags"]).all() and (h["roots"] == e["roots"]).all(), f"replay {rep}"
    for ty, p1, r in zip(five_types, p1s, g.results):
        o = oracle_for(oracle, ty["system"], ty["n"], ty["p0"], p1, ty["start"])
        h = r.to_host()
        compare(hcx, oracle, h, o, f"graph type n={ty['n']} K={ty['K']}", exact_flags=False)
        counts_equal(hcx, h, o, f"graph type n={ty['n']}")
    g.close()


def test_graph_survives_larger_eager_launch(hcx_mod, five_types):
    """A graph owns its launch workspace: an eager launch of the same
    problems with MORE instances (which needs a larger workspace) between two
    replays leaves the replays equal to the eager result of the captured
    batch."""
    import torch
    from hcx import workloads
    hcx = hcx_mod
    tys = [five_types[0], five_types[2]]
    probs = [hcx.Problem.from_system(t["system"], t["start"], p0=t["p0"]) for t in tys]
    small = [hcx.Batch(p, 3, target_params=workloads.targets(t["k"], 3, seed=5 + t["n"])) for p, t in zip(probs, tys)]
    big = [hcx.Batch(p, 40, target_params=workloads.targets(t["k"], 40, seed=9 + t["n"])) for p, t in zip(probs, tys)]
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    ref = [r.to_host() for r in hcx.track_multi(probs, small, opts)]
    torch.cuda.synchronize()
    g = hcx.Graph(probs, small, opts)
    g.launch()
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    for i in range(2):
        hcx.track_multi(probs, big, opts, stream=stream)        # larger eager launch, another stream
        hcx.track_multi(probs, big, opts)                        # and the default stream
        torch.cuda.synchronize()
        for r in g.results:
            r.roots.fill_(float("nan"))
            r.flags.fill_(0)
        g.launch()
        torch.cuda.synchronize()
        for e, r in zip(ref, g.results):
            h = r.to_host()
            assert (h["flags"] == e["flags"]).all() and (h["roots"] == e["roots"]).all(), f"round {i}"
    g.close()
    hcx.stream_release(stream)


# ------------------------------------------------------------------ six served types
CONFIG5 = [(3, 8), (6, 64), (9, 142), (12, 583), (16, 40), (30, 312)]


@pytest.fixture(scope="module")
def six_types(hcx_mod, gpu):
    from hcx import workloads
    print("[six types] seeding start roots", flush=True)
    types = workloads.quadratic_types(CONFIG5, seed=5)
    probs = [hcx_mod.Problem.from_system(t["system"], t["start"], p0=t["p0"]) for t in types]
    return types, probs


def wait(server, rid, tid, timeout=120.0):
    t0 = time.time()
    while True:
        r = server.result(rid, tid)
        if r is not None:
            return r
        assert time.time() - t0 < timeout, "request did not finish"
        time.sleep(0.001)


def test_six_served_types_vs_oracle(hcx_mod, oracle_mod, six_types):
    from hcx import workloads
    hcx, oracle = hcx_mod, oracle_mod
    types, probs = six_types
    srv = hcx.Server(hcx.default_opts(precision=hcx.PREC_MIXED), hbm_budget_bytes=512 << 20)
    tids = [srv.add_type(p) for p in probs]
    xs = [workloads.targets(ty["k"], 1, 5000 + ti) for ti, ty in enumerate(types)]
    rids = [srv.submit(tids[ti], xs[ti][0]) for ti in range(len(types))]
    results = [wait(srv, rid, tids[ti]) for ti, rid in enumerate(rids)]
    for ti, ty in enumerate(types):
        info, roots, flags = results[ti]
        t0 = time.time()
        o = oracle_for(oracle, ty["system"], ty["n"], ty["p0"], xs[ti], ty["start"])
        g = {"flags": flags, "roots": roots}
        print(f"[served n={ty['n']} K={ty['K']}] oracle {time.time() - t0:.1f} s", flush=True)
        compare(hcx, oracle, g, o, f"served n={ty['n']}", exact_flags=False)
        counts_equal(hcx, g, o, f"served n={ty['n']}")
        assert info["refined"] == int(((flags & hcx.FLAG_REFINED) != 0).sum())
    srv.close()


@pytest.mark.parametrize("prec", [1, 0])
def test_large_types_alone_equal_in_mix(hcx_mod, six_types, prec):
    from hcx import workloads
    hcx = hcx_mod
    types, probs = six_types
    srv = hcx.Server(hcx.default_opts(precision=prec), hbm_budget_bytes=512 << 20, max_batch_requests=16)
    tids = [srv.add_type(p) for p in probs]
    probe = {ti: workloads.targets(types[ti]["k"], 1, 700 + ti)[0] for ti in (3, 4, 5)}  # n = 12, 16, 30
    alone = {ti: wait(srv, srv.submit(tids[ti], x), tids[ti]) for ti, x in probe.items()}
    rng = np.random.default_rng(3)
    ids, mid = [], {}
    for i in range(48):
        t = int(rng.integers(0, len(types)))
        ids.append((srv.submit(tids[t], workloads.targets(types[t]["k"], 1, 9000 + i)[0]), t))
        if i in (10, 20, 30):
            ti = {10: 3, 20: 4, 30: 5}[i]
            mid[ti] = srv.submit(tids[ti], probe[ti])
    srv.drain()
    for ti, rid in mid.items():
        m = wait(srv, rid, tids[ti])
        assert (alone[ti][2] == m[2]).all() and np.array_equal(alone[ti][1], m[1]), f"type n={types[ti]['n']}"
    for rid, t in ids:
        assert wait(srv, rid, t) is not None
    st = srv.stats()
    print("serving stats", st, flush=True)
    assert st["completed"] == st["submitted"]
    srv.close()


# ------------------------------------------------------------------ device collection
def test_device_collect_equals_oracle(hcx_mod, oracle_mod, gpu):
    """hcx_collect on the device gives the oracle's counts and flags bit for
    bit, including the reference's unique_ref (Find_Unique_Sols control
    flow) and real_ref, on hand-built duplicate patterns and tracked roots."""
    import torch
    from test_collect_ref import CASES, roots_from
    hcx, oracle = hcx_mod, oracle_mod
    rng = np.random.default_rng(11)
    labels, flags = [], []
    K = 8
    for c, (_, lab, fl, _) in enumerate(CASES):
        # each case is one instance, padded with distinct tracked roots
        pad = K - len(lab)
        labels += [f"{c}:{x}" for x in lab] + [f"{c}:pad{q}" for q in range(pad)]
        for f in list(fl) + [0x01] * pad:
            if f & 0x01:
                f |= 0x04 | (0x200 if rng.random() < 0.5 else 0)
            flags.append(f)
    I = len(CASES)
    roots = roots_from(labels, n=3)
    f = np.array(flags, dtype=np.uint32)
    r = hcx.Result(I * K, 3, I)
    r.roots.copy_(torch.from_numpy(roots))
    r.flags.copy_(torch.from_numpy(f.view(np.int32)))
    gc = hcx.collect(r, hcx.default_opts())
    gf = r.flags.cpu().numpy().view(np.uint32)
    oc, of = oracle.collect(3, I, K, roots, f)
    assert gc == oc, (gc, oc)
    assert (gf == of).all()
    assert gc["unique_ref"] == sum(c[3] for c in CASES) + sum(K - len(c[1]) for c in CASES)


# ------------------------------------------------------------------ group LU through linear systems
def linear_system_problem(hcx, A, b0, b1):
    """H(x, t) = A x - (b0 + t (b1 - b0)): the Jacobian is A at every point,
    so every predictor and corrector solve is an LU of A."""
    n = A.shape[0]
    exps, eq = [], []
    for i in range(n):
        for j in range(n):
            e = np.zeros(n, int)
            e[j] = 1
            exps.append(e)
            eq.append(i)
        exps.append(np.zeros(n, int))
        eq.append(i)
    s = hcx.System.from_monomials(np.array(exps), np.array(eq))
    hxi, hti = s.tables()
    Cn = n * (n + 1)
    hx = np.zeros((Cn + 1, 2), complex)
    ht = np.zeros((Cn + 1, 1), complex)
    for i in range(n):
        for j in range(n):
            hx[i * (n + 1) + j, 0] = A[i, j]
        c = i * (n + 1) + n
        hx[c] = [-b0[i], -(b1[i] - b0[i])]
        ht[c, 0] = -(b1[i] - b0[i])
    hx[Cn, 0] = 1
    return s, hxi, hti, Cn, hx, ht


@pytest.mark.parametrize("case", ["ties", "zero_column", "random12"])
def test_group_lu_through_linear_systems(hcx_mod, oracle_mod, gpu, case):
    """FP64: flags and step counts equal the oracle's; a non-singular system
    lands on A^-1 b1 = the oracle LU's solution.  'ties': equal cabs1 in the
    pivot column (first maximum in current row order wins, as the reference's
    strict > scan, dev-cgesv-batched-small.cuh:58-64); 'zero_column': a zero
    pivot column is skipped and the back substitution yields NaN, so the
    path fails the same way in both."""
    import torch
    hcx, oracle = hcx_mod, oracle_mod
    rng = np.random.default_rng(5)
    if case == "ties":
        A = np.array([[1, 2, 0, 1j], [1j, 1, 1, 0], [-1, 0, 3, 2], [0.5 + 0.5j, 1, 1j, 1]], dtype=complex)
    elif case == "zero_column":
        A = np.array([[1, 0, 2], [3, 0, 1], [1, 0, 1]], dtype=complex)
    else:
        A = rng.normal(size=(12, 12)) + 1j * rng.normal(size=(12, 12))
    n = A.shape[0]
    x0 = rng.normal(size=n) + 1j * rng.normal(size=n)
    b0 = A @ x0
    b1 = rng.normal(size=n) + 1j * rng.normal(size=n)
    s, hxi, hti, Cn, hx, ht = linear_system_problem(hcx, A, b0, b1)
    prob = hcx.Problem(n, Cn, hxi, hti, x0[None])
    b = hcx.Batch(prob, 1, coeffs_hx=hx[None], coeffs_ht=ht[None])
    g = hcx.track(prob, b, hcx.default_opts(precision=hcx.PREC_FP64)).to_host()
    torch.cuda.synchronize()
    op = oracle.OracleProblem(n, Cn, hxi, hti, 1)
    o = op.track(x0[None], hx[None], ht[None], threads=1)
    print(f"[lu {case}] gpu flags {g['flags'][0]:#x} steps {g['steps'][0]}; oracle {o['flags'][0]:#x} "
          f"steps {o['steps'][0]}", flush=True)
    assert g["flags"][0] == o["flags"][0]
    assert g["steps"][0] == o["steps"][0] and g["newton"][0] == o["newton"][0]
    if case == "zero_column":
        assert not (g["flags"][0] & hcx.FLAG_TRACKED)
        return
    xs, _, info = oracle.lu_solve(A, b1)
    assert info == 0
    assert g["flags"][0] & hcx.FLAG_REFINED
    assert rel_err(g["roots"][0], xs) <= 1e-12 * np.linalg.cond(A)
    assert rel_err(g["roots"][0], np.linalg.solve(A, b1)) <= 1e-12 * np.linalg.cond(A)
