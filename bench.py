"""Benchmark: converged paths/s + p99 solve latency of the MI355X tracker,
mixed-precision batched (one persistent launch per batch) versus the plain
FP64 per-problem launch, on the same GPU, same systems, parameters and seeds.

Workload (default): n=6 parameter homotopy, 12 monomials per equation, k=24
parameters; 64 start roots from a seed total-degree solve at p0; 1024 target
instances -> 65,536 paths per step.  A step tracks every path of the batch and
collects the counts (refined / real / unique).

Prints ONE JSON line (rank 0).  `value` = converged (FP64-refined) paths per
second of the mixed batched form with all inputs resident in HBM.
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


def percentile(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a, dtype=float), q)) if len(a) else float("nan")


def cpu_baseline(sysf, p0, start, p1, seconds):
    """Oracle (FP64 C++ restatement) on host cores over a bounded sample."""
    import numpy as np
    import oracle
    threads = os.cpu_count() or 1
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
    t0 = time.perf_counter()
    out = op.track(start, hx, ht, threads=threads)
    el = time.perf_counter() - t0
    conv = int(((out["flags"] & oracle.FLAG_REFINED) != 0).sum())
    return {"value": conv / el, "unit": "converged paths/s", "cores": threads, "kind": "port",
            "sample": f"{n_inst} instances x {len(start)} paths ({n_inst * len(start)} paths) of the same workload, "
                      f"FP64 oracle, {threads} threads, {el:.1f} s",
            "paths_per_s": n_inst * len(start) / el}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--lanes", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

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
    opts_mixed = hcx.default_opts(precision=hcx.PREC_MIXED, lanes_per_path=args.lanes)
    opts_fp64 = hcx.default_opts(precision=hcx.PREC_FP64, lanes_per_path=args.lanes)
    batch = hcx.Batch(prob, I, target_params=p1)
    res = hcx.Result(P, prob.n, I)
    stream = torch.cuda.current_stream()

    def solve(opts, result):
        hcx.track(prob, batch, opts, result=result, stream=stream)
        return hcx.collect(result, opts, stream=stream)

    # ------------------------------------------------------------ mixed batched
    log(f"bench: mixed batched, {I} instances x {K} paths, warmup {args.warmup}, steps {args.steps}")
    for _ in range(args.warmup):
        solve(opts_mixed, res)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    counts = None
    for _ in range(args.steps):
        counts = solve(opts_mixed, res)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = counts["refined"] / (elapsed / args.steps)
    h = res.to_host()
    lat_ms = (h["inst_done"][:I] - h["batch_start"][0]) / DEVICE_CLOCK_HZ * 1e3
    log(f"bench: mixed {ms_per_step:.2f} ms/step, {value:.4g} converged paths/s, counts {counts}")

    # kernel-only timing + FLOPs for the roofline (coefficient tables precomputed)
    hxs = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[0] for i in range(I)])
    hts = np.stack([sysf.param_homotopy_coeffs(p0, p1[i])[1] for i in range(I)])
    kbatch = hcx.Batch(prob, I, coeffs_hx=hxs, coeffs_ht=hts)
    kres = hcx.Result(P, prob.n, I)
    hcx.track(prob, kbatch, opts_mixed, result=kres, stream=stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = max(3, args.steps)
    ev[0].record(stream)
    for _ in range(reps):
        hcx.track(prob, kbatch, opts_mixed, result=kres, stream=stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    kernel_ms = ev[0].elapsed_time(ev[1]) / reps
    kh = kres.to_host()
    flops = float(hcx.flops_per_path(info, kh["steps"], kh["newton"], kh["refine"]).sum())
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    log(f"bench: track kernel {kernel_ms:.3f} ms, {flops:.4g} algorithmic FLOP -> {achieved:.2f} TFLOP/s")

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
                  "refined_mixed": int(okm.sum()), "refined_fp64": int(ok6.sum()),
                  "real_mixed": int(((h["flags"][:P] & hcx.FLAG_REAL) != 0).sum()),
                  "real_fp64": int(((flags64 & hcx.FLAG_REAL) != 0).sum()),
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
        "dtype": "c64+c128 (FP32 predictor/Jacobian, FP64 endgame corrector + refinement)",
        "data": "synthetic (seeded CN(0,1) parameters)",
        "config": {"workload": "n=6 parameter homotopy, 12 monomials/eq, k=24 params, 64 start roots, "
                               f"{I} target instances ({P} paths per step)",
                   "instances": I, "paths_per_instance": K, "seed": args.seed,
                   "lanes_per_path": args.lanes or prob.n},
        "p50_ms": percentile(lat_ms, 50),
        "p99_ms": percentile(lat_ms, 99),
        "latency_note": "per-instance completion time from batch launch (device clock)",
        "counts": counts,
        "fp64_per_problem": fp64,
        "speedup_vs_fp64_per_problem": (value / fp64["value"]) if fp64 else None,
        "parity_mixed_vs_fp64": parity,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TFLOPS, "traffic": profiled_traffic(),
                     "traffic_unit": "bytes/launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
                     "kernel_ms": kernel_ms, "algorithmic_flop": flops,
                     "note": "LAWN-41 FLOPs of the reference algorithm from per-path device step/Newton "
                             "counters; peak = vector FP64 (= non-packed FP32 rate)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
