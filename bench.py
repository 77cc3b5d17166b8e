"""Benchmark: converged paths/s + p99 solve latency of the MI355X tracker,
batched (one persistent launch per batch) versus the plain FP64 per-problem
launch, on the same GPU, same systems, parameters and seeds.

Workload (default): n=6 parameter homotopy, 12 monomials per equation, k=24
parameters; 64 start roots from a seed total-degree solve at p0; 1024 target
instances -> 65,536 paths per step.  A step tracks every path of the batch and
collects the counts (refined / real / unique).

Prints ONE JSON line (rank 0).  `value` = converged (FP64-refined) paths per
second of the batched form (FP64 micro-step tracker by default; the mixed
FP32/FP64 tracker is measured beside it) with all inputs resident in HBM.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "homotopy-continuation-tracker-on-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "converged paths/sec + p99 solve latency, mixed/batched vs plain FP64, same GPU"
PEAK_VALU_TFLOPS = 78.6   # MI355X vector FP64 (= non-packed FP32 issue rate)
PEAK_FP32_PACKED = 157.3  # vector FP32 with packed math (not used by these kernels)
DEVICE_CLOCK_HZ = 100e6   # s_memrealtime


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic():
    """HBM bytes per launch of the track kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    import re
    path = os.path.join(ROOT, "profiles", "track_kernel_traffic_n6_param_homotopy_1024x64.md")
    try:
        m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def flops_split(info, h, precision_mixed):
    """Algorithmic FLOPs (LAWN 41) of a solve's paths from the per-path device
    counters, split by the precision they ran in: in mixed precision the RK4
    stages of a path run in FP32 unless the path was re-tracked in FP64
    (HCX_FLAG_ESCALATED), and the corrector, refinement and coefficient
    evaluation at t = 1 run in FP64."""
    import numpy as np
    import hcx
    steps = h["steps"].astype(np.float64)
    newton = h["newton"].astype(np.float64)
    refine = h["refine"].astype(np.float64)
    stage = 4.0 * steps * info["flops_stage"]
    rest = (newton + refine) * info["flops_newton"] + (refine > 0) * info["flops_coef"]
    if not precision_mixed:
        return 0.0, float((stage + rest).sum())
    esc = (h["flags"] & hcx.FLAG_ESCALATED) != 0
    f32 = float(stage[~esc].sum())
    f64 = float(stage[esc].sum() + rest.sum())
    return f32, f64


def roofline_valu(f32, f64, seconds, kernel_ms=None):
    """Vector-FLOP roofline of the solve: achieved TFLOP/s over the timed
    steady state against MI355X's published vector peaks for the same
    FP32/FP64 instruction mix (157.3 FP32, 78.6 FP64 TFLOP/s)."""
    total = f32 + f64
    achieved = total / seconds / 1e12
    peak = total / (f32 / PEAK_FP32_PACKED + f64 / PEAK_VALU_TFLOPS) if total > 0 else PEAK_VALU_TFLOPS
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "algorithmic_flop": total, "flop_fp32": f32, "flop_fp64": f64,
            "achieved_fp32": f32 / seconds / 1e12, "achieved_fp64": f64 / seconds / 1e12,
            "frac_vs_unpacked_peak": achieved / PEAK_VALU_TFLOPS,
            "kernel_ms": kernel_ms,
            "note": "LAWN-41 FLOPs of the reference algorithm from per-path device step/Newton counters over the "
                    "timed steady state (solves in flight included); peak = published vector peaks weighted by "
                    "the FP32/FP64 mix; FP32 runs unpacked here (78.6 TFLOP/s issue bound, frac_vs_unpacked_peak)"}


def percentile(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a, dtype=float), q)) if len(a) else float("nan")


def host_threads():
    """Threads the CPU baseline may use: the cores this process may run on
    (sched_getaffinity), at most 16 -- the host share of one GPU on the
    benchmark machines, whatever the machine reports in total."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail)), avail, os.cpu_count() or 1


def _cpu_rate(op, start, hx, ht, threads, seconds):
    import numpy as np
    import oracle
    conv = paths = reps = 0
    n_inst = hx.shape[0]
    t0 = time.perf_counter()
    while True:
        out = op.track(start, hx, ht, threads=threads)
        conv += int(((out["flags"] & oracle.FLAG_REFINED) != 0).sum())
        paths += n_inst * len(start)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            return conv / el, paths / el, reps, paths, el


def cpu_baseline(sysf, p0, start, p1, seconds):
    """Oracle (FP64 C++ restatement of the reference's tracker) on host cores
    over a bounded sample of the same workload: all usable threads (the
    headline value) and one thread."""
    import numpy as np
    import oracle
    threads, avail, total = host_threads()
    hxi, hti = sysf.tables()
    op = oracle.OracleProblem(sysf.info["num_vars"], sysf.info["num_coeffs"], hxi, hti, 1)
    A = sysf.affine()
    # calibrate on one instance
    hx, ht = oracle.p2c_affine(A, p0, p1[:1])
    t0 = time.perf_counter()
    op.track(start, hx, ht, threads=threads)
    t1 = time.perf_counter() - t0
    n_inst = int(max(1, min(len(p1), seconds / max(t1, 1e-4))))
    hx, ht = oracle.p2c_affine(A, p0, p1[:n_inst])
    # repeat the sample until it holds about `seconds` of CPU work
    value, pps, reps, paths, el = _cpu_rate(op, start, hx, ht, threads, seconds)
    # one thread on a smaller sample of the same instances
    n1 = max(1, n_inst // max(1, threads))
    v1, pps1, reps1, paths1, el1 = _cpu_rate(op, start, hx[:n1], ht[:n1], 1, seconds * 0.5)
    return {"value": value, "unit": "converged paths/s", "cores": threads, "kind": "port",
            "sample": f"{n_inst} instances x {len(start)} paths of the same workload, tracked {reps} times "
                      f"({paths} paths), FP64 oracle, {threads} threads, {el:.1f} s",
            "paths_per_s": pps,
            "single_thread": {"value": v1, "unit": "converged paths/s", "cores": 1,
                              "sample": f"{n1} instances x {len(start)} paths, tracked {reps1} times "
                                        f"({paths1} paths), {el1:.1f} s"},
            "host_cpus": {"usable": avail, "reported": total, "used": threads}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="fp64",
                    help="precision of the batched solve of the headline (the other one is measured "
                         "beside it), of the per-type sweep and of the serving leg")
    ap.add_argument("--inflight", type=int, default=4,
                    help="configs 1, 2: solves in flight on separate streams (the duplicate re-track "
                         "and tail of one solve overlap the next ones)")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="config 2: persistent grid density (waves per CU; 0: the kernel's occupancy; "
                         "< 0: occupancy less that many)")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="default run: seconds of each serving window (fixed rate, then the offered 2000 req/s)")
    ap.add_argument("--serve-rate", type=float, default=800.0,
                    help="default run: the fixed Poisson request rate the latency is stated at")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-refprob", action="store_true",
                    help="skip the reference's problems one instance per launch in the headline run")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="end with one 256 MiB device copy, the reference dispatch for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE scale factors")
    ap.add_argument("--config", type=int, default=2,
                    help="workload: 2 (default, the headline line), 1, 3, 4, 5, 6 (per-type cost "
                         "sweep of the serving types) or 7 (the reference's own problems, one instance "
                         "per launch, beside the reference's published GPU times)")
    ap.add_argument("--problems", default="",
                    help="config 7: comma-separated problem names (default: all with a published time)")
    ap.add_argument("--replicas", type=int, default=0,
                    help="config 7: replica lane groups per path (0 = the library's choice, 1 = off)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.config != 2:
        return {1: bench_config1, 3: bench_config3, 4: bench_config4, 5: bench_config5,
                6: bench_type_sweep, 7: bench_reference_problems}[args.config](args)

    import numpy as np
    import torch

    import hcx
    from hcx import workloads

    assert hcx.device_count() > 0, "no HIP device"
    torch.cuda.init()
    dev = torch.device("cuda:0")

    # ------------------------------------------------------------ workload
    sysf, p0 = workloads.config2_system(args.seed)
    log("bench: seeding start roots (total-degree solve at p0)")
    start = workloads.seed_start_roots(sysf, p0, 64, seed=args.seed)
    I = args.instances
    p1 = workloads.targets(24, I, seed=args.seed * 7 + 1)
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    info = prob.info
    K = prob.num_tracks
    P = I * K
    opts_head = hcx.default_opts(precision=hcx.PREC_MIXED if args.precision == "mixed" else hcx.PREC_FP64,
                                  lanes_per_path=args.lanes, waves_per_cu=args.waves_per_cu)
    opts_fp64 = hcx.default_opts(precision=hcx.PREC_FP64, lanes_per_path=args.lanes)
    batch = hcx.Batch(prob, I, target_params=p1)
    res = hcx.Result(P, prob.n, I)
    stream = torch.cuda.current_stream()

    # solves in flight: one resident problem; each solve has its own batch
    # (coefficient workspace), result and stream (launch workspaces belong to
    # streams); the host enqueues solve i before it waits for the counts of
    # solve i - 1
    nf = max(1, args.inflight)
    lanes = [(prob, batch, res, stream)]
    for _ in range(nf - 1):
        lanes.append((prob, hcx.Batch(prob, I, target_params=p1), hcx.Result(P, prob.n, I), torch.cuda.Stream()))

    def run_steps(opts, count):
        counts = None
        for i in range(count):
            pk, bk, rk, sk = lanes[i % nf]
            hcx.track(pk, bk, opts, result=rk, stream=sk)
            if i + 1 >= nf:
                pj, bj, rj, sj = lanes[(i + 1 - nf) % nf]
                counts = hcx.collect(rj, opts, stream=sj)
        for i in range(max(0, count - nf + 1), count):
            pj, bj, rj, sj = lanes[i % nf]
            counts = hcx.collect(rj, opts, stream=sj)
        return counts

    # ------------------------------------------------------------ batched (headline)
    head = "mixed" if args.precision == "mixed" else "fp64"
    other = "fp64" if head == "mixed" else "mixed"
    opts_other = hcx.default_opts(precision=hcx.PREC_FP64 if other == "fp64" else hcx.PREC_MIXED,
                                  lanes_per_path=args.lanes, waves_per_cu=args.waves_per_cu)
    # clock ramp: untimed solves for about half a second, so that the timed
    # steps do not start on an idle GPU's clocks (reported as prewarm_s)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.5:
        run_steps(opts_head, nf)
        torch.cuda.synchronize()
    prewarm_s = time.perf_counter() - t0
    log(f"bench: {head} batched, {I} instances x {K} paths, warmup {args.warmup}, steps {args.steps}, "
        f"{nf} solves in flight (after {prewarm_s:.2f} s of untimed solves)")
    run_steps(opts_head, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counts = run_steps(opts_head, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = counts["refined"] / (elapsed / args.steps)
    h = res.to_host()
    lat_ms = (h["inst_done"][:I] - h["batch_start"][0]) / DEVICE_CLOCK_HZ * 1e3
    f32, f64 = flops_split(info, h, args.precision == "mixed")
    log(f"bench: {head} {ms_per_step:.2f} ms/step, {value:.4g} converged paths/s, counts {counts}")

    # the same batched loop in the other precision: reported beside the
    # headline, not instead of it
    run_steps(opts_other, args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    c_o = run_steps(opts_other, args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    other_batched = {"value": c_o["refined"] / (el / args.steps), "unit": "converged paths/s",
                     "ms_per_step": el * 1e3 / args.steps, "counts": c_o,
                     "counts_identical": bool(c_o == counts)}
    log(f"bench: {other} batched {other_batched['ms_per_step']:.2f} ms/step, {other_batched['value']:.4g} paths/s")

    # kernel-only timing + FLOPs for the roofline (coefficient tables precomputed)
    hxs = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[0] for i in range(I)])
    hts = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[1] for i in range(I)])
    kbatch = hcx.Batch(prob, I, coeffs_hx=hxs, coeffs_ht=hts)
    kres = hcx.Result(P, prob.n, I)
    hcx.track(prob, kbatch, opts_head, result=kres, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = max(3, args.steps)
    ev[0].record(stream)
    for _ in range(reps):
        hcx.track(prob, kbatch, opts_head, result=kres, stream=stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    kernel_ms = ev[0].elapsed_time(ev[1]) / reps
    roof = roofline_valu(f32, f64, ms_per_step * 1e-3, kernel_ms=kernel_ms)
    roof["traffic"] = profiled_traffic()
    roof["traffic_unit"] = "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)"
    log(f"bench: one solve alone {kernel_ms:.3f} ms; steady state {roof['algorithmic_flop']:.4g} FLOP per step "
        f"(FP32 {f32:.3g}, FP64 {f64:.3g}) -> {roof['achieved']:.2f} TFLOP/s, frac {roof['frac']:.4f}")

    # ------------------------------------------------------------ plain FP64 per-problem
    fp64 = None
    if not args.no_fp64:
        log("bench: plain FP64 per-problem launches")
        sres = hcx.Result(K, prob.n, 1)
        sb = [hcx.Batch(prob, 1, target_params=p1[i:i + 1]) for i in range(I)]
        for i in range(min(8, I)):
            hcx.track(prob, sb[i], opts_fp64, result=sres, static=True, stream=stream)
        torch.cuda.synchronize()
        lat = []
        conv = 0
        roots64 = np.zeros((P, prob.n), dtype=np.complex128)
        flags64 = np.zeros(P, dtype=np.uint32)
        t0 = time.perf_counter()
        for i in range(I):
            ti = time.perf_counter()
            hcx.track(prob, sb[i], opts_fp64, result=sres, static=True, stream=stream)
            c = hcx.collect(sres, opts_fp64, stream=stream)
            lat.append((time.perf_counter() - ti) * 1e3)
            conv += c["refined"]
            hh = sres.to_host()
            roots64[i * K:(i + 1) * K] = hh["roots"]
            flags64[i * K:(i + 1) * K] = hh["flags"]
        el64 = time.perf_counter() - t0 - 0.0
        # the host copies above are bookkeeping for the parity check; time them out
        t0 = time.perf_counter()
        for i in range(min(64, I)):
            sres.to_host()
        copy_ms = (time.perf_counter() - t0) / min(64, I) * 1e3
        el64 -= copy_ms * 1e-3 * I
        v64 = conv / el64
        okm = (h["flags"][:P] & hcx.FLAG_REFINED) != 0
        ok6 = (flags64 & hcx.FLAG_REFINED) != 0
        both = okm & ok6
        err = (np.abs(h["roots"][both] - roots64[both]).max(axis=1) /
               np.maximum(1, np.abs(roots64[both]).max(axis=1)))
        fp64 = {"value": v64, "unit": "converged paths/s", "converged": int(conv),
                "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                "launches": I, "seconds": el64}
        parity = {"counts_identical": bool(int(okm.sum()) == int(ok6.sum())),
                  "refined_batched": int(okm.sum()), "refined_fp64_per_problem": int(ok6.sum()),
                  "real_batched": int(((h["flags"][:P] & hcx.FLAG_REAL) != 0).sum()),
                  "real_fp64_per_problem": int(((flags64 & hcx.FLAG_REAL) != 0).sum()),
                  "max_root_rel_err": float(err.max()) if len(err) else 0.0}
        log(f"bench: fp64 per-problem {v64:.4g} converged paths/s, p50 {fp64['p50_ms']:.3f} ms "
            f"p99 {fp64['p99_ms']:.3f} ms; parity {parity}")
    else:
        parity = None

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle)")
        cpu = cpu_baseline(sysf, p0, start, p1, args.cpu_seconds)
        log(f"bench: cpu {cpu}")

    serving = None
    if not args.no_serve:
        serving = serving_leg(args)

    refprob = None
    if not args.no_refprob:
        log("bench: the reference's problems, one instance per launch")
        rows = reference_problem_rows(args, 5)
        refprob = {"faster_than_published": sum(r["fp64_per_problem"]["ms_p50"] < r["reference_gpu_hc_ms"]
                                                for r in rows),
                   "problems": len(rows),
                   "ms": {r["problem"]: [round(r["fp64_per_problem"]["ms_p50"], 3), r["reference_gpu_hc_ms"]]
                          for r in rows},
                   "note": "[hcx FP64 ms p50, reference GPU-HC ms]: one target instance per launch (P2C + "
                           "tracker, the reference's settings) against the reference's published times "
                           "(README.md:45-63, complex FP32, GPU not stated); bench.py --config 7 for details"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "converged paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64+c128 (FP32 RK4 predictor and Jacobian, FP64 corrector, duplicate re-track and refinement)"
                  if args.precision == "mixed" else
                  "c128 (FP64 throughout: RK4 predictor, Jacobian, LU, Newton corrector, refinement)"),
        "data": "synthetic (seeded CN(0,1) parameters)",
        "config": {"workload": "n=6 parameter homotopy, 12 monomials/eq, k=24 params, 64 start roots, "
                               f"{I} target instances ({P} paths per step)",
                   "instances": I, "paths_per_instance": K, "seed": args.seed,
                   "lanes_per_path": args.lanes or prob.n, "solves_in_flight": nf,
                   "waves_per_cu": args.waves_per_cu},
        "p50_ms": serving["fixed_rate"]["p50_ms"] if serving else None,
        "p99_ms": serving["fixed_rate"]["p99_ms"] if serving else None,
        "latency_note": (f"solve latency (submit -> result on the host) of the serving loop at a fixed "
                         f"{args.serve_rate:.0f} req/s Poisson stream over the six problem types; see serving"),
        "serving": serving,
        "batch_instance_completion_ms": {"p50": percentile(lat_ms, 50), "p99": percentile(lat_ms, 99),
                                         "note": "per-instance completion time from batch launch (device clock)"},
        "counts": counts,
        "fp64_per_problem": fp64,
        "speedup_vs_fp64_per_problem": (value / fp64["value"]) if fp64 else None,
        f"{other}_batched": other_batched,
        "prewarm_s": prewarm_s,
        "parity_batched_vs_fp64_per_problem": parity,
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_problems_one_instance": refprob,
    }
    print(json.dumps(out))
    if args.pmc_calibrate:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()


def _roofline(flops, kernel_ms):
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_VALU_TFLOPS, "traffic": None, "kernel_ms": kernel_ms, "algorithmic_flop": flops}


def _rel_err(a, b):
    import numpy as np
    return np.abs(a - b).max(axis=1) / np.maximum(1.0, np.abs(b).max(axis=1))


def bench_config1(args):
    """Dense quadratic n=3 total-degree homotopy: single-instance latency (8
    paths) of the FP64 per-problem launch and of the mixed launch, and
    throughput over R instances of the same support with independent seeds."""
    import numpy as np
    import torch
    import hcx
    import oracle
    from hcx import workloads
    c = workloads.config1(args.seed)
    td = c["homotopy"]
    hxi, hti = td.tables()
    nc = td.info["num_coeffs"]
    prob = hcx.Problem(3, nc, hxi, hti, c["start"])
    stream = torch.cuda.current_stream()
    out = {"metric": METRIC, "config": {"workload": "config 1: dense quadratic n=3 total-degree homotopy, 8 paths"}}
    for name, prec, static in (("fp64_per_problem", hcx.PREC_FP64, True), ("mixed", hcx.PREC_MIXED, False)):
        opts = hcx.default_opts(precision=prec)
        b = hcx.Batch(prob, 1, coeffs_hx=c["coeffs_hx"], coeffs_ht=c["coeffs_ht"])
        r = hcx.Result(8, 3, 1)
        lat = []
        for i in range(args.warmup + 50):
            t0 = time.perf_counter()
            hcx.track(prob, b, opts, result=r, static=static, stream=stream)
            torch.cuda.synchronize()
            if i >= args.warmup:
                lat.append((time.perf_counter() - t0) * 1e3)
        h = r.to_host()
        out[name] = {"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                     "refined": int(((h["flags"] & hcx.FLAG_REFINED) != 0).sum())}
    ref = oracle.OracleProblem(3, nc, hxi, hti, 1).track(c["start"], c["coeffs_hx"], c["coeffs_ht"], threads=1)
    out["oracle_refined"] = int(((ref["flags"] & oracle.FLAG_REFINED) != 0).sum())
    out["max_root_rel_err_vs_oracle"] = float(_rel_err(h["roots"], ref["roots"]).max())
    # throughput: R instances of the dense support, independent coefficients and gamma
    R = 10000
    hxs, hts = [], []
    for i in range(R):
        sysf = hcx.System.synthetic(3, 0, num_quadratic=3, monomials=0, seed=1000 + i)
        _, hx, ht = sysf.total_degree(sysf.eval_coeffs(None), hcx.random_gamma(5000 + i))
        hxs.append(hx)
        hts.append(ht)
    hxs, hts = np.stack(hxs), np.stack(hts)
    # solves in flight on separate streams (one resident problem; each solve
    # has its own batch and result, launch workspaces belong to the streams)
    nf = max(1, args.inflight)
    lanes = [(prob, hcx.Batch(prob, R, coeffs_hx=hxs, coeffs_ht=hts), hcx.Result(R * 8, 3, R), stream)]
    for _ in range(nf - 1):
        lanes.append((prob, hcx.Batch(prob, R, coeffs_hx=hxs, coeffs_ht=hts), hcx.Result(R * 8, 3, R),
                      torch.cuda.Stream()))
    opts = hcx.default_opts(precision=hcx.PREC_MIXED)
    for pk, bk, rk, sk in lanes:
        hcx.track(pk, bk, opts, result=rk, stream=sk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        pk, bk, rk, sk = lanes[i % nf]
        hcx.track(pk, bk, opts, result=rk, stream=sk)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / args.steps
    cnt = hcx.collect(lanes[0][2], opts, stream=stream)
    out.update({"value": cnt["refined"] / el, "unit": "converged paths/s", "ms_per_step": el * 1e3,
                "instances": R, "solves_in_flight": nf, "counts": cnt})
    print(json.dumps(out))


def bench_config3(args):
    """Uneven batch: five problem types mixed over 4096 instances, one
    persistent multi-type launch captured as a HIP graph, autotuned shapes."""
    import numpy as np
    import torch
    import hcx
    from hcx import workloads
    I = 4096 if args.instances == 1024 else args.instances
    log("config3: seeding start roots of five types")
    types = workloads.quadratic_types(workloads.CONFIG3_TYPES, seed=args.seed)
    kinds = workloads.mixed_instances(len(types), I, args.seed)
    opts = hcx.default_opts(precision=_prec(args))
    probs, batches, tunes = [], [], []
    for ti, ty in enumerate(types):
        cnt = int((kinds == ti).sum())
        prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
        p1 = workloads.targets(ty["k"], cnt, seed=args.seed * 31 + ti)
        b = hcx.Batch(prob, cnt, target_params=p1)
        sample = hcx.Batch(prob, min(cnt, max(8, 4096 // ty["K"])), target_params=p1[:max(8, 4096 // ty["K"])])
        tunes.append(hcx.autotune(prob, sample, opts))
        probs.append(prob)
        batches.append(b)
    log(f"config3: autotuned {[(t['lanes_per_path'], t['specialize'], t['waves_per_cu']) for t in tunes]}")
    g = hcx.Graph(probs, batches, opts)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        g.launch(stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(args.steps):
        g.launch(stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / args.steps
    counts = [hcx.collect(r, opts, stream=stream) for r in g.results]
    conv = sum(c["refined"] for c in counts)
    paths = sum(b.I * p.num_tracks for p, b in zip(probs, batches))
    flops = 0.0
    lat = []
    start = g.results[0].to_host()["batch_start"][0]
    for p, r in zip(probs, g.results):
        h = r.to_host()
        flops += float(hcx.flops_per_path(p.info, h["steps"], h["newton"], h["refine"]).sum())
        lat += list((h["inst_done"][: r.I] - start) / DEVICE_CLOCK_HZ * 1e3)
    # the same batch as separate per-type persistent launches, for comparison
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for p, b in zip(probs, batches):
        hcx.track(p, b, opts, stream=stream)
    torch.cuda.synchronize()
    sep_ms = (time.perf_counter() - t0) * 1e3
    out = {"metric": METRIC, "value": conv / (ms * 1e-3), "unit": "converged paths/s", "ms_per_step": ms,
           "steps": args.steps, "warmup": args.warmup, "dtype": "c64+c128 mixed" if args.precision == "mixed" else "c128",
           "config": {"workload": f"config 3: uneven batch of 5 quadratic types (n,K)={workloads.CONFIG3_TYPES}, "
                                  f"{I} instances mixed uniformly, {paths} paths, one graph-captured launch",
                      "instances_per_type": [b.I for b in batches]},
           "paths": paths, "converged": conv, "counts_per_type": counts,
           "autotune": tunes, "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
           "separate_launches_ms": sep_ms, "roofline": _roofline(flops, ms)}
    print(json.dumps(out))


def bench_config4(args):
    """Mixed-precision tracking at n=12 (6 quadratic + 6 cubic, 20 monomials,
    k=36), 512 start roots x 2048 instances; per-root error of the mixed form
    against the FP64 form on every path and against the oracle on a sample."""
    import numpy as np
    import torch
    import hcx
    import oracle
    from hcx import workloads
    I = 2048 if args.instances == 1024 else args.instances
    sysf, p0 = workloads.config4_system(args.seed)
    log("config4: seeding 512 start roots")
    start = workloads.seed_start_roots(sysf, p0, 512, seed=args.seed)
    p1 = workloads.targets(36, I, seed=args.seed * 11 + 3)
    prob = hcx.Problem.from_system(sysf, start, p0=p0)
    batch = hcx.Batch(prob, I, target_params=p1)
    stream = torch.cuda.current_stream()
    out = {"metric": METRIC, "config": {"workload": f"config 4: n=12 (6 quadratic + 6 cubic, 20 monomials, k=36), "
                                                    f"512 start roots x {I} instances = {512 * I} paths"}}
    res = {}
    for name, prec in (("mixed", hcx.PREC_MIXED), ("fp64", hcx.PREC_FP64)):
        opts = hcx.default_opts(precision=prec)
        r = hcx.Result(I * 512, 12, I)
        hcx.track(prob, batch, opts, result=r, stream=stream)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(args.steps):
            hcx.track(prob, batch, opts, result=r, stream=stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / args.steps
        h = r.to_host()
        cnt = hcx.collect(r, opts, stream=stream)
        flops = float(hcx.flops_per_path(prob.info, h["steps"], h["newton"], h["refine"]).sum())
        res[name] = h
        out[name] = {"ms_per_launch": ms, "converged_paths_per_s": cnt["refined"] / (ms * 1e-3), "counts": cnt,
                     "escalated": int(((h["flags"] & hcx.FLAG_ESCALATED) != 0).sum()),
                     "roofline": _roofline(flops, ms)}
        log(f"config4: {name} {ms:.2f} ms, counts {cnt}")
    okm = (res["mixed"]["flags"] & hcx.FLAG_REFINED) != 0
    ok6 = (res["fp64"]["flags"] & hcx.FLAG_REFINED) != 0
    both = okm & ok6
    err = _rel_err(res["mixed"]["roots"][both], res["fp64"]["roots"][both])
    out["mixed_vs_fp64"] = {"counts_identical": bool(okm.sum() == ok6.sum()), "refined": [int(okm.sum()), int(ok6.sum())],
                            "max_root_rel_err": float(err.max()), "median_root_rel_err": float(np.median(err))}
    # oracle on a bounded sample of instances
    ns = 4
    hxi, hti = sysf.tables()
    op = oracle.OracleProblem(12, sysf.info["num_coeffs"], hxi, hti, 1)
    ohx, oht = oracle.p2c_affine(sysf.affine(), p0, p1[:ns])
    t0 = time.perf_counter()
    o = op.track(start, ohx, oht, threads=min(16, os.cpu_count() or 1))
    el = time.perf_counter() - t0
    okr = (o["flags"] & oracle.FLAG_REFINED) != 0
    sm = okm[: ns * 512]
    b2 = sm & okr
    e2 = _rel_err(res["mixed"]["roots"][: ns * 512][b2], o["roots"][b2])
    out["mixed_vs_oracle_sample"] = {"instances": ns, "paths": ns * 512, "refined_gpu": int(sm.sum()),
                                     "refined_oracle": int(okr.sum()), "max_root_rel_err": float(e2.max()),
                                     "oracle_seconds": el, "oracle_threads": min(16, os.cpu_count() or 1)}
    out["value"] = out["mixed"]["converged_paths_per_s"]
    out["unit"] = "converged paths/s"
    out["speedup_mixed_vs_fp64_batched"] = out["fp64"]["ms_per_launch"] / out["mixed"]["ms_per_launch"]
    print(json.dumps(out))


def _serve(srv, tids, types, rate, duration, seed):
    """Poisson arrivals at `rate` for `duration` seconds; returns the stats."""
    import numpy as np
    from hcx import workloads
    rng = np.random.default_rng(seed)
    srv.reset_stats()
    n_req = int(rate * duration)
    gaps = rng.exponential(1.0 / rate, n_req)
    kinds = rng.integers(0, len(types), n_req)
    params = [workloads.targets(types[k]["k"], 1, seed * 1000003 + i)[0] for i, k in enumerate(kinds)]
    pending = []
    t_start = time.perf_counter()
    t_next = t_start
    late = 0.0
    log(f"config5: {n_req} requests at {rate:.0f}/s for {duration:.0f} s")
    last_log = t_start
    for i in range(n_req):
        t_next += gaps[i]
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            if pending and t_next - now > 0.0005:
                rid, k = pending[0]
                if srv.result(rid, k) is not None:
                    pending.pop(0)
                    continue
            if t_next - now > 0.0003:
                time.sleep(0.0001)
        late = max(late, time.perf_counter() - t_next)
        pending.append((srv.submit(tids[kinds[i]], params[i]), int(kinds[i])))
        if time.perf_counter() - last_log > 5:
            last_log = time.perf_counter()
            st = srv.stats()
            log(f"config5: t={last_log - t_start:.1f}s submitted {st['submitted']} completed {st['completed']} "
                f"inflight {st['inflight_requests']} p99 {st['p99_ms']:.2f} ms")
    t_sub = time.perf_counter() - t_start
    srv.drain(600.0)
    t_end = time.perf_counter() - t_start
    for rid, k in pending:
        srv.result(rid, k)
    st = srv.stats()
    return {"offered_requests_per_s": rate, "seconds": duration, "requests": n_req,
            "requests_per_s": st["completed"] / t_end, "converged_paths_per_s": st["converged_paths"] / t_end,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"], "max_ms": st["max_ms"],
            "submit_seconds": t_sub, "total_seconds": t_end, "max_submit_lateness_ms": late * 1e3, "stats": st}


def _prec(args):
    import hcx
    return hcx.PREC_MIXED if args.precision == "mixed" else hcx.PREC_FP64


def serving_leg(args):
    """Bounded serving windows for the default run: Poisson arrivals over the
    six resident types n in {3, 6, 9, 12, 16, 30} at a fixed rate (the stated
    p50 / p99 solve latency) and at the offered 2000 req/s."""
    import hcx
    from hcx import workloads
    log("bench: serving leg, seeding start roots of six types")
    types = workloads.quadratic_types(workloads.CONFIG5_TYPES, seed=args.seed)
    probs = [hcx.Problem.from_system(t["system"], t["start"], p0=t["p0"]) for t in types]
    srv = hcx.Server(hcx.default_opts(precision=_prec(args)), hbm_budget_bytes=4 << 30, max_batch_requests=64,
                     streams=4)
    tids = [srv.add_type(p) for p in probs]
    for ti, ty in enumerate(types):
        srv.submit(tids[ti], workloads.targets(ty["k"], 1, 90000 + ti)[0])
    srv.drain()
    out = {"types": [[t["n"], t["K"]] for t in types], "hbm_budget_bytes": 4 << 30, "streams": 4,
           "max_batch_requests": 64}
    for key, rate, seed in (("fixed_rate", args.serve_rate, args.seed + 1), ("offered_2000", 2000.0, args.seed)):
        r = _serve(srv, tids, types, rate, args.serve_seconds, seed=seed)
        out[key] = {"offered_rps": rate, "served_rps": r["requests_per_s"], "p50_ms": r["p50_ms"],
                    "p99_ms": r["p99_ms"], "mean_ms": r["mean_ms"], "max_ms": r["max_ms"],
                    "converged_paths_per_s": r["converged_paths_per_s"], "requests": r["requests"],
                    "seconds": r["total_seconds"]}
        log(f"bench: serving {key}: {out[key]}")
    srv.close()
    for p in probs:
        p.close()
    return out


def _fp64_per_problem_mix(probs, types, per_type, seed):
    """The plain FP64 per-problem launch (one static launch per request, the
    reference's launch shape) over the same type mix: converged paths/s with
    equally many requests of every type."""
    import torch
    import hcx
    from hcx import workloads
    o = hcx.default_opts(precision=hcx.PREC_FP64)
    conv = 0
    secs = 0.0
    per = []
    for ti, (p, ty) in enumerate(zip(probs, types)):
        res = hcx.Result(p.num_tracks, p.n, 1)
        tg = workloads.targets(ty["k"], per_type + 1, seed + 7 * ti)
        b0 = hcx.Batch(p, 1, target_params=tg[:1])
        hcx.track(p, b0, o, result=res, static=True)
        torch.cuda.synchronize()
        batches = [hcx.Batch(p, 1, target_params=tg[i + 1:i + 2]) for i in range(per_type)]
        c = 0
        t0 = time.perf_counter()
        for b in batches:
            hcx.track(p, b, o, result=res, static=True)
            c += hcx.collect(res, o)["refined"]
        el = time.perf_counter() - t0
        conv += c
        secs += el
        per.append({"n": ty["n"], "K": ty["K"], "ms_per_request": el / per_type * 1e3})
    return {"converged_paths_per_s": conv / secs, "requests_per_type": per_type, "per_type": per}


def bench_config5(args):
    """Serving loop: Poisson arrivals of single-instance requests over six
    resident quadratic types n in {3,6,9,12,16,30}, continuous batching with
    admission by HBM budget.  Runs at the offered 2000 req/s (throughput at
    saturation) and at a fixed sustainable rate (latency), and compares the
    served converged paths/s with the plain FP64 per-problem launch over the
    same mix."""
    import hcx
    from hcx import workloads
    rate = float(os.environ.get("HCX_SERVE_RATE", "2000"))
    duration = float(os.environ.get("HCX_SERVE_SECONDS", "30"))
    fixed = float(os.environ.get("HCX_SERVE_FIXED_RATE", "400"))
    fixed_hi = float(os.environ.get("HCX_SERVE_FIXED_RATE_HIGH", "800"))
    log("config5: seeding start roots of six types")
    types = workloads.quadratic_types(workloads.CONFIG5_TYPES, seed=args.seed)
    probs = [hcx.Problem.from_system(t["system"], t["start"], p0=t["p0"]) for t in types]
    opts = hcx.default_opts(precision=_prec(args))
    srv = hcx.Server(opts, hbm_budget_bytes=4 << 30, max_batch_requests=64, streams=4)
    tids = [srv.add_type(p) for p in probs]
    for ti, ty in enumerate(types):
        srv.submit(tids[ti], workloads.targets(ty["k"], 1, 90000 + ti)[0])
    srv.drain()
    saturated = _serve(srv, tids, types, rate, duration, seed=args.seed)
    steady = _serve(srv, tids, types, fixed, duration, seed=args.seed + 1)
    steady_hi = _serve(srv, tids, types, fixed_hi, duration, seed=args.seed + 2) if fixed_hi > 0 else None
    srv.close()
    log("config5: plain FP64 per-problem launches over the same mix")
    fp64 = _fp64_per_problem_mix(probs, types, per_type=24, seed=args.seed * 31)
    out = {"metric": METRIC, "value": saturated["converged_paths_per_s"], "unit": "converged paths/s",
           "config": {"workload": f"config 5: serving loop, Poisson arrivals, types (n,K)={workloads.CONFIG5_TYPES} "
                                  f"drawn uniformly, one instance per request, {args.precision}",
                      "hbm_budget_bytes": 4 << 30, "streams": 4, "max_batch_requests": 64},
           "at_offered_rate": saturated, "at_fixed_rate": steady, "at_fixed_rate_high": steady_hi,
           "p50_ms": steady["p50_ms"], "p99_ms": steady["p99_ms"], "latency_rate_req_s": fixed,
           "fp64_per_problem": fp64,
           "speedup_vs_fp64_per_problem": saturated["converged_paths_per_s"] / fp64["converged_paths_per_s"]}
    print(json.dumps(out))


def bench_type_sweep(args):
    """GPU cost per request of each serving type: the six quadratic types
    n in {3, 6, 9, 12, 16, 30}, one batched mixed-precision launch of
    `--instances` requests (default 64) per type, timed over `--steps`
    launches.  Also the workload of the rocprofv3 counter profiles."""
    import torch
    import hcx
    from hcx import workloads
    B = 64 if args.instances == 1024 else args.instances
    types = workloads.quadratic_types(workloads.CONFIG5_TYPES, seed=5)
    opts = hcx.default_opts(precision=_prec(args))
    stream = torch.cuda.current_stream()
    rows = []
    for ty in types:
        prob = hcx.Problem.from_system(ty["system"], ty["start"], p0=ty["p0"])
        b = hcx.Batch(prob, B, target_params=workloads.targets(ty["k"], B, seed=77))
        r = hcx.Result(B * prob.num_tracks, prob.n, B)
        hcx.track(prob, b, opts, result=r, stream=stream)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(args.steps):
            hcx.track(prob, b, opts, result=r, stream=stream)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / args.steps
        cnt = hcx.collect(r, opts, stream=stream)
        h = r.to_host()
        rows.append({"n": ty["n"], "K": ty["K"], "requests": B, "ms_per_launch": ms, "ms_per_request": ms / B,
                     "converged_paths_per_s": cnt["refined"] / (ms * 1e-3), "refined": cnt["refined"],
                     "steps_mean": float(h["steps"].mean()), "steps_max": int(h["steps"].max())})
        log(f"sweep: n={ty['n']} K={ty['K']}: {ms:.2f} ms per {B} requests")
        del b, r
        prob.close()
    print(json.dumps({"metric": METRIC, "config": {"workload": f"per-type cost sweep, {B} requests per type, "
                                                               f"{args.precision}, one batched launch each"},
                      "types": rows}))



# The reference's published GPU-HC times (README.md:45-63: one launch tracking
# all paths of one problem instance, complex FP32, GPU not stated), for the
# problems whose committed fixture holds the full start-solution set.
REFERENCE_README_MS = {
    "uncalibrated_trifocal_rel_pos_CY": ("trifocal relative pose, unknown focal length", 668, 16.67),
    "generalized_3views_4pts": ("generalized 3-view relative pose from 4 points", 583, 4.77),
    "5pt_rel_pos_full_form_quat": ("5-point relative pose with depths", 40, 7.26),
    "optimal_PnP_quat": ("optimal PnP (quaternion)", 80, 0.77),
    "dual_reciever_TDOA_5pt": ("dual-receiver TDOA 5-points", 24, 3.17),
    "distorted_2view_triangulation": ("distorted 2-view triangulation", 28, 0.83),
    "5pt_rel_pos_alg_form_quat": ("5-point relative pose (reduced/algebraic form)", 40, 1.77),
    "optimal_P4P_abs_pos": ("optimal P4P absolute pose", 32, 0.43),
    "6pt_RS_abs_pos": ("6-point rolling shutter absolute pose", 20, 0.65),
    "3pt_rel_pos_homog": ("3-point relative pose with homography constraint", 8, 0.84),
    "PnP_unkn_principal_pt": ("PnP with unknown principal point", 8, 3.45),
    "rel_pos_quiver": ("relative pose using quiver, unknown focal length", 20, 0.75),
    "P3P": ("P3P, absolute pose", 8, 0.15),
}


def bench_reference_problems(args):
    """The reference's own problems (tables, coefficient maps, start
    solutions and settings from the committed fixtures), one instance per
    launch in the reference's launch shape (static grid, one solve at a
    time, synchronous), FP64, timed with HIP events around P2C + tracker;
    beside each, the reference's published GPU-HC time for the same problem
    and path count (complex FP32, GPU not stated)."""
    rows = reference_problem_rows(args, max(5, args.steps))
    print(json.dumps({"metric": METRIC, "config": {"workload": "the reference's problems, one instance per "
                                                               "launch, reference settings, FP64"},
                      "problems": rows,
                      "note": "reference times: README.md:45-63, complex FP32 on an unstated GPU"}))


def reference_problem_rows(args, reps):
    import numpy as np
    import torch
    import hcx
    fix = os.path.join(ROOT, "tests", "golden", "reference")
    stream = torch.cuda.current_stream()
    rows = []
    only = set(filter(None, args.problems.split(",")))
    for name, (title, paths, ref_ms) in REFERENCE_README_MS.items():
        if only and name not in only:
            continue
        d = np.load(os.path.join(fix, name + ".npz"), allow_pickle=False)
        start = d["start_sols"]
        if start.shape[0] != paths:
            continue
        if str(d["kind"]) != "P2C":
            continue
        n = int(d["n"])
        hx4 = d["hx"].reshape(n, n, int(d["hx_terms"]), int(d["hx_parts"]))
        ht3 = d["ht"].reshape(n, int(d["ht_terms"]), int(d["ht_parts"]))
        prob = hcx.Problem(n, int(d["C"]), hx4, ht3, start, max_order_t=int(d["M"]),
                           cmap=(d["cmap_scalar"], d["cmap_params"]), p0=d["start_params"])
        k = int(d["k"])
        tp = d["target_params"] if bool(d["target_complete"]) else hcx.random_cn(args.seed + 1, k)
        b = hcx.Batch(prob, 1, target_params=tp.reshape(1, k))
        out = {}
        for label, static, prec in (("fp64_per_problem", True, hcx.PREC_FP64),
                                    ("fp64_persistent", False, hcx.PREC_FP64)):
            o = hcx.default_opts(precision=prec, max_steps=int(d["max_steps"]), max_corr=int(d["max_corr"]),
                                 incr_steps=int(d["incr_steps"]), replicas=args.replicas)
            r = hcx.track(prob, b, o, static=static, stream=stream)
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(stream)
                hcx.track(prob, b, o, result=r, static=static, stream=stream)
                ev[1].record(stream)
                torch.cuda.synchronize()
                ts.append(ev[0].elapsed_time(ev[1]))
            h = r.to_host()
            out[label] = {"ms_p50": percentile(ts, 50), "ms_min": float(min(ts)),
                          "tracked": int(((h["flags"] & hcx.FLAG_TRACKED) != 0).sum()),
                          "refined": int(((h["flags"] & hcx.FLAG_REFINED) != 0).sum()),
                          "steps_max": int(h["steps"].max())}
        row = {"problem": name, "title": title, "n": int(d["n"]), "paths": paths,
               "reference_gpu_hc_ms": ref_ms, **out,
               "speedup_vs_reference": ref_ms / out["fp64_per_problem"]["ms_p50"]}
        rows.append(row)
        log(f"reference problem {name}: n={row['n']} paths={paths}: {out['fp64_per_problem']['ms_p50']:.3f} ms "
            f"(persistent {out['fp64_persistent']['ms_p50']:.3f} ms) vs reference {ref_ms} ms")
        del b
        prob.close()
    return rows


if __name__ == "__main__":
    main()
