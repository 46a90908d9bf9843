"""Benchmark: batched GJK pair-queries/s on one MI355X.

Workload: 1,048,576 mixed-primitive pairs (sphere / capsule / cylinder / box /
ellipsoid / convex hull with log-uniform 8..4096 vertices), seed 1, scene and
pairs resident in HBM before the timed region.  A step is one pass of the
batched GJK distance kernel over all pairs.

Prints ONE JSON line with throughput, the roofline of the GJK kernel
(algorithmic bytes per launch / average launch time, against 8 TB/s), the
parity of the last step against the FP64 oracle, and the oracle timed on the
host cores as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GJK+EPA pair-queries/s at ≤1e-5 rel err vs FP64; p50/p99 request latency"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--pairs", type=int, default=1 << 20)
    p.add_argument("--seed", type=int, default=1)
    p.add_argument("--lanes", type=str, default=None,
                   help="lanes per pair for the 8 work classes, comma separated")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-parity", action="store_true", help="skip the oracle parity check")
    p.add_argument("--cpu-runs", type=int, default=3)
    return p.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the collision path shards by pair range "
              "with no exchange, so a multi-GPU figure is N times this one and is not "
              "measured here (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from distance3d_amd import kernels, workloads

    t0 = time.time()
    w = workloads.mixed_pairs(seed=args.seed, n_pairs=args.pairs)
    log("workload generated in %.1fs: %d pairs, %d hull vertices"
        % (time.time() - t0, w.n_pairs, w.scene.n_vertices))
    dev = w.scene.to_device("cuda:0")
    pa = torch.from_numpy(w.pair_a).cuda()
    pb = torch.from_numpy(w.pair_b).cuda()
    n = w.n_pairs
    out = kernels.GjkResult(
        torch.empty(n, dtype=torch.float64, device="cuda"),
        torch.empty((n, 3), dtype=torch.float64, device="cuda"),
        torch.empty((n, 3), dtype=torch.float64, device="cuda"),
        torch.empty(n, dtype=torch.int32, device="cuda"),
        torch.empty(n, dtype=torch.int32, device="cuda"), None, None)

    lanes = [int(x) for x in args.lanes.split(",")] if args.lanes else None
    ws = torch.empty(kernels.gjk_workspace_bytes(n), dtype=torch.uint8, device="cuda")

    def step():
        kernels.gjk_distance_batch(dev, pa, pb, lanes_per_pair=lanes, out=out, workspace=ws)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("warmup done")

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        step()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    kernel_ms = np.array([a.elapsed_time(b) for a, b in ev])
    ms_per_step = t_total * 1e3 / args.steps
    value = n / (ms_per_step / 1e3)
    log("timed %d steps: %.3f ms/step, kernel avg %.3f ms" % (args.steps, ms_per_step,
                                                               kernel_ms.mean()))

    # algorithmic bytes per pair: 12 (n_A + n_B) FP32 hull bytes, 2 x 64 pose and
    # parameters, 8 pair ids, 60 outputs (dist, two closest points, status)
    bytes_per_launch = float(np.sum(w.hull_bytes_per_pair()) + n * (128 + 8 + 60))
    achieved = bytes_per_launch / (kernel_ms.mean() / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None, "kernel": "gjk_distance_kernel",
                "kernel_ms": round(float(kernel_ms.mean()), 4),
                "bytes_per_launch": bytes_per_launch}

    iters = out.iterations.cpu().numpy()
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "pair-queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 simplex / f32 hull sweeps",
        "data": "synthetic",
        "config": {"workload": "mixed_pairs (1,048,576 mixed-primitive GJK pairs)",
                   "pairs": n, "seed": args.seed, "hull_vertices": "log-uniform 8..4096",
                   "query": "GJK distance + closest points", "epa": False,
                   "lanes_per_pair": lanes or list(__import__("distance3d_amd._native",
                                                              fromlist=["x"]).DEFAULT_LANES),
                   "mean_iterations": round(float(iters.mean()), 3)},
        "roofline": roofline,
    }

    if not args.no_parity or not args.no_cpu:
        from oracle import oracle
    if not args.no_parity:
        t0 = time.time()
        o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b)
        dist = out.dist.cpu().numpy()
        sep = (o["dist"] > 1e-9) & (o["dist"] < np.finfo(float).max)
        rel = np.abs(dist[sep] - o["dist"][sep]) / o["dist"][sep]
        band = (o["dist"] < 1e-9) | (dist < 1e-9)
        flag_mismatch = int(np.sum(((dist == 0.0) != (o["dist"] == 0.0)) & ~band))
        result["parity"] = {
            "vs": "FP64 oracle", "pairs": n,
            "max_rel_dist_err": float(rel.max(initial=0.0)),
            "p99_rel_dist_err": float(np.percentile(rel, 99)) if len(rel) else 0.0,
            "bit_identical_frac": float(np.mean(dist == o["dist"])),
            "flag_mismatches_outside_1e-9_band": flag_mismatch,
            "intersecting_frac": float(np.mean(o["dist"] == 0.0)),
        }
        log("parity checked in %.1fs" % (time.time() - t0))
    if not args.no_cpu:
        cores = oracle.num_threads()
        oracle.gjk_distance(w.scene, w.pair_a[:65536], w.pair_b[:65536])  # warm-up
        times = []
        for _ in range(args.cpu_runs):
            t0 = time.perf_counter()
            oracle.gjk_distance(w.scene, w.pair_a, w.pair_b)
            times.append(time.perf_counter() - t0)
        cpu_rate = n / float(np.median(times))
        result["cpu_baseline"] = {
            "value": round(cpu_rate, 1), "unit": "pair-queries/s", "cores": cores,
            "kind": "port",
            "sample": "all %d pairs of the same workload, median of %d runs (%.1fs total), "
                      "C FP64 restatement with OpenMP" % (n, args.cpu_runs, sum(times))}
        result["speedup_vs_cpu"] = round(value / cpu_rate, 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
