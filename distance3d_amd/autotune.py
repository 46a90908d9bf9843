"""Launch tuning for the batched GJK: lanes per pair for each work class.

The GJK pipeline sorts pairs into 8 classes by their total hull vertex count
(0, <=32, <=128, <=512, <=1024, <=2048, <=4096, >4096).  Two tables are kept in
``tuning.json`` beside this file:

* ``lanes_per_pair`` (throughput, :func:`tune`): for each class, the GJK
  kernels are timed on that class's own pairs (taken from a large sample
  workload) with every candidate lane count and the fastest is kept;
* ``serving_lanes_per_pair`` (latency, :func:`tune_latency`): coordinate
  descent over the classes, minimising the median time of one GJK + EPA
  batch of a single large request on an otherwise idle GPU -- more lanes per
  hull pair shorten each support sweep, which is what bounds a small batch.

:func:`tuned_lanes` returns either table.

Usage: ``python -m distance3d_amd.autotune [--pairs N] [--serving-requests M]``
on the GPU.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _native

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CLASS_EDGES = (0, 32, 128, 512, 1024, 2048, 4096)
CANDIDATES = {0: (1, 2, 4), 1: (1, 2, 4, 8), 2: (2, 4, 8, 16), 3: (4, 8, 16, 32),
              4: (8, 16, 32, 64), 5: (8, 16, 32, 64), 6: (16, 32, 64), 7: (16, 32, 64)}
LATENCY_CANDIDATES = {0: (1, 2, 4), 1: (4, 8, 16, 32), 2: (8, 16, 32, 64), 3: (16, 32, 64),
                      4: (16, 32, 64), 5: (32, 64), 6: (64,), 7: (64,)}


def work_class(nverts):
    """Host restatement of the device work-class rule (d3d.h)."""
    nverts = np.asarray(nverts)
    cls = np.full(nverts.shape, 7, dtype=np.int32)
    for c in range(6, -1, -1):
        cls[nverts <= CLASS_EDGES[c]] = c
    cls[nverts == 0] = 0
    return cls


def pair_classes(scene, pair_a, pair_b):
    i = scene.shape_i32
    nv = np.where(i[:, 0] == _native.HULL, i[:, 1], 0).astype(np.int64)
    return work_class(nv[pair_a] + nv[pair_b])


def tuned_lanes(path=TUNING_PATH, kind="throughput"):
    """The tuned table (``kind`` "throughput" or "latency"), or None when the
    tuning file or that table is missing."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    lanes = t.get("lanes_per_pair" if kind == "throughput" else "serving_lanes_per_pair")
    return None if lanes is None else [int(x) for x in lanes]


def _time(dev, pa, pb, lanes, reps):
    import torch
    from . import kernels
    n = pa.numel()
    out = kernels.GjkResult(torch.empty(n, dtype=torch.float64, device="cuda"),
                            torch.empty((n, 3), dtype=torch.float64, device="cuda"),
                            torch.empty((n, 3), dtype=torch.float64, device="cuda"),
                            torch.empty(n, dtype=torch.int32, device="cuda"), None, None, None)
    ws = torch.empty(kernels.gjk_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    kernels.gjk_distance_batch(dev, pa, pb, lanes_per_pair=lanes, out=out, workspace=ws)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        kernels.gjk_distance_batch(dev, pa, pb, lanes_per_pair=lanes, out=out, workspace=ws)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(workload, reps=3, log=print):
    import torch
    dev = workload.scene.to_device()
    cls = pair_classes(workload.scene, workload.pair_a, workload.pair_b)
    lanes = list(_native.DEFAULT_LANES)
    report = {}
    for c in range(0, 8):
        idx = np.nonzero(cls == c)[0]
        if len(idx) < 256:
            continue
        pa = torch.from_numpy(workload.pair_a[idx]).cuda()
        pb = torch.from_numpy(workload.pair_b[idx]).cuda()
        times = {}
        for g in CANDIDATES[c]:
            trial = list(lanes)
            trial[c] = g
            times[g] = _time(dev, pa, pb, trial, reps)
        best = min(times, key=times.get)
        lanes[c] = best
        report[c] = {"pairs": int(len(idx)), "ms": {str(k): round(v * 1e3, 4) for k, v in
                                                    times.items()}, "best": best}
        log("class %d (%d pairs): %s -> %d lanes" % (c, len(idx), report[c]["ms"], best))
    return lanes, report


def combined_step_ms(workload, lanes, reps=5):
    """Best time of the whole GJK + EPA step over every pair of the workload
    (all classes in one launch, as bench.py runs it)."""
    import torch
    from . import kernels
    dev = workload.scene.to_device()
    pa = torch.from_numpy(workload.pair_a).cuda()
    pb = torch.from_numpy(workload.pair_b).cuda()
    n = pa.numel()
    out = kernels.GjkEpaResult.empty(n, "cuda")
    ws = torch.empty(kernels.gjk_epa_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    kernels.gjk_epa_batch(dev, pa, pb, out=out, workspace=ws, lanes_per_pair=lanes)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        kernels.gjk_epa_batch(dev, pa, pb, out=out, workspace=ws, lanes_per_pair=lanes)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return 1e3 * best


def tune_latency(scene, poses, requests, start=None, sweeps=2, reps=3, log=print):
    """Lanes per class minimising the median GJK + EPA time of one request
    batch on an idle GPU.  ``requests``: serving.Request objects over
    ``scene`` (a HostScene) and the (P, 12) ``poses`` table."""
    import torch
    from . import kernels
    dev = scene.to_device()
    P = torch.from_numpy(np.asarray(poses, dtype=np.float64).reshape(-1, 12)).cuda()
    batches = []
    for r in requests:
        n = r.n_pairs
        t = [torch.from_numpy(x).cuda() for x in (r.pair_a, r.pair_b, r.pose_a, r.pose_b)]
        view, pa, pb = kernels.instance_pairs(dev, t[0], t[1], P, P, t[2], t[3], P.shape[0])
        out = kernels.GjkEpaResult.empty(n, "cuda")
        ws = torch.empty(kernels.gjk_epa_workspace_bytes(n), dtype=torch.uint8, device="cuda")
        batches.append((view, pa, pb, out, ws))

    def measure(lanes):
        best = []
        for view, pa, pb, out, ws in batches:
            ts = []
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                kernels.gjk_epa_batch(view, pa, pb, out=out, workspace=ws, lanes_per_pair=lanes)
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            best.append(min(ts))
        return float(np.median(best))

    lanes = list(start or _native.DEFAULT_LANES)
    measure(lanes)  # warm-up
    report = {"start": {"lanes": list(lanes), "ms": round(1e3 * measure(lanes), 4)}}
    for sweep in range(sweeps):
        for c in range(8):
            times = {}
            for g in LATENCY_CANDIDATES[c]:
                trial = list(lanes)
                trial[c] = g
                times[g] = measure(trial)
            lanes[c] = min(times, key=times.get)
            report["sweep%d_class%d" % (sweep, c)] = {str(k): round(v * 1e3, 4) for k, v in
                                                      times.items()}
            log("latency sweep %d class %d: %s -> %d lanes" % (
                sweep, c, {k: round(v * 1e3, 3) for k, v in times.items()}, lanes[c]))
    report["final"] = {"lanes": list(lanes), "ms": round(1e3 * measure(lanes), 4)}
    return lanes, report


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--pairs", type=int, default=1 << 20)
    p.add_argument("--serving-requests", type=int, default=48,
                   help="large requests (>= 2048 pairs) of the serving workload to tune on")
    p.add_argument("--out", default=TUNING_PATH)
    args = p.parse_args()
    from . import workloads
    w = workloads.mixed_pairs(seed=1, n_pairs=args.pairs)
    lanes, report = tune(w)
    # classes overlap in the combined launch, so the per-class optimum is
    # checked against the installed table on the whole step; the faster stays
    installed = tuned_lanes()
    if installed is not None and installed != lanes:
        t_new, t_old = combined_step_ms(w, lanes), combined_step_ms(w, installed)
        report["combined_ms"] = {"tuned": [lanes, round(t_new, 4)],
                                 "installed": [installed, round(t_old, 4)]}
        print("combined step: tuned %s %.3f ms, installed %s %.3f ms" % (lanes, t_new, installed, t_old))
        if t_old < t_new:
            lanes = installed
    scene, poses, reqs = workloads.serving_workload(n_requests=20000)
    big = [r for r in reqs if r.n_pairs >= 2048][:args.serving_requests]
    slanes, sreport = tune_latency(scene, poses, big, start=lanes)
    with open(args.out, "w") as f:
        json.dump({"lanes_per_pair": lanes, "workload": "mixed_pairs seed 1, %d pairs" % args.pairs,
                   "classes": report,
                   "serving_lanes_per_pair": slanes,
                   "serving_workload": "%d requests of >= 2048 pairs from the serving stream "
                                       "(seed 4), one batch each on an idle GPU" % len(big),
                   "serving": sreport}, f, indent=1)
    print(json.dumps({"lanes_per_pair": lanes, "serving_lanes_per_pair": slanes}))


if __name__ == "__main__":
    main()
