"""Launch tuning for the batched GJK: lanes per pair for each work class and
the launch shape (the tile of pairs a wavefront claims, blocks per compute
unit).

The GJK pipeline sorts pairs into 16 work classes: 8 size classes by their
total hull vertex count (0, <=32, <=128, <=512, <=1024, <=2048, <=4096,
>4096), each split into pairs with one hull and pairs of two hulls (class + 8).
The tables kept in ``tuning.json`` beside this file:

* ``lanes_per_pair`` (throughput, :func:`tune`): for each class, the GJK
  kernels are timed on that class's own pairs (taken from a large sample
  workload) with every candidate lane count and the fastest is kept;
* ``launch`` (throughput, :func:`tune_launch`): coordinate descent over the
  launch shape -- ``blocks_per_cu`` persistent GJK blocks per compute unit,
  ``claim_factor`` (a wavefront claims claim_factor * 64 / G pairs of a class
  at a time: the tile of pairs it works through), ``epa_blocks_per_cu``
  persistent EPA wavefronts per compute unit -- minimising the whole GJK + EPA
  step;
* ``serving_lanes_per_pair`` (latency, :func:`tune_latency`): coordinate
  descent over the classes, minimising the median time of one GJK + EPA
  batch of a single large request on an otherwise idle GPU -- more lanes per
  hull pair shorten each support sweep, which is what bounds a small batch.

:func:`tuned_lanes` and :func:`tuned_launch` return the tables.  The selection
(:func:`pick`, :func:`descend`) is plain host logic over measured times.

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
NUM_CLASSES = _native.NUM_CLASSES
_SIZE_CANDIDATES = {0: (1, 2, 4), 1: (1, 2, 4, 8), 2: (2, 4, 8, 16), 3: (4, 8, 16, 32),
                    4: (8, 16, 32, 64), 5: (8, 16, 32, 64), 6: (16, 32, 64), 7: (16, 32, 64)}
_SIZE_LATENCY = {0: (1, 2, 4), 1: (4, 8, 16, 32), 2: (8, 16, 32, 64), 3: (16, 32, 64),
                 4: (16, 32, 64), 5: (32, 64), 6: (64,), 7: (64,)}
# the same candidates for pairs of one hull (class c) and of two (c + 8)
CANDIDATES = {c: _SIZE_CANDIDATES[c % 8] for c in range(NUM_CLASSES)}
LATENCY_CANDIDATES = {c: _SIZE_LATENCY[c % 8] for c in range(NUM_CLASSES)}


LAUNCH_CANDIDATES = {"blocks_per_cu": (2, 3), "claim_factor": (1, 2, 4),
                     "epa_blocks_per_cu": (4, 8, 12)}
LAUNCH_DEFAULT = {"blocks_per_cu": 3, "claim_factor": 2, "epa_blocks_per_cu": 8}


def pick(times):
    """The candidate with the smallest time; on equal times the one listed
    first (``times`` is an ordered mapping candidate -> time)."""
    best = None
    for k, t in times.items():
        if best is None or t < times[best]:
            best = k
    return best


def descend(measure, start, candidates, sweeps=1, log=None):
    """Coordinate descent: for each key of ``candidates`` in order, every
    candidate value is measured with the other keys held at their current
    values and the fastest is kept; ``sweeps`` passes over the keys.
    ``measure(config) -> time``.  Returns (config, report)."""
    cfg = dict(start)
    report = {"start": {"config": dict(cfg), "time": measure(dict(cfg))}}
    for sw in range(sweeps):
        for key, values in candidates.items():
            times = {}
            for v in values:
                trial = dict(cfg)
                trial[key] = v
                times[v] = measure(trial)
            cfg[key] = pick(times)
            report["sweep%d_%s" % (sw, key)] = {str(v): t for v, t in times.items()}
            if log:
                log("%s: %s -> %s" % (key, times, cfg[key]))
    report["final"] = {"config": dict(cfg), "time": measure(dict(cfg))}
    return cfg, report


def work_class(nverts):
    """Host restatement of the device size-class rule (d3d.h)."""
    nverts = np.asarray(nverts)
    cls = np.full(nverts.shape, 7, dtype=np.int32)
    for c in range(6, -1, -1):
        cls[nverts <= CLASS_EDGES[c]] = c
    cls[nverts == 0] = 0
    return cls


def pair_classes(scene, pair_a, pair_b):
    """Work class of each pair: size class, + 8 when both colliders are hulls."""
    i = scene.shape_i32
    nv = np.where(i[:, 0] == _native.HULL, i[:, 1], 0).astype(np.int64)
    na, nb = nv[pair_a], nv[pair_b]
    return work_class(na + nb) + np.where((na > 0) & (nb > 0), 8, 0).astype(np.int32)


def tuned_lanes(path=TUNING_PATH, kind="throughput"):
    """The tuned table (``kind`` "throughput" or "latency"), or None when the
    tuning file or that table is missing."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    lanes = t.get("lanes_per_pair" if kind == "throughput" else "serving_lanes_per_pair")
    return None if lanes is None else _native.expand_lanes(lanes)


def tuned_launch(path=TUNING_PATH, kind="throughput"):
    """The tuned launch shape (dict for ``launch=``), or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    launch = t.get("launch" if kind == "throughput" else "serving_launch")
    return None if launch is None else {k: int(v) for k, v in launch.items()}


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
    for c in range(NUM_CLASSES):
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
        best = pick(times)
        lanes[c] = best
        report[c] = {"pairs": int(len(idx)), "ms": {str(k): round(v * 1e3, 4) for k, v in
                                                    times.items()}, "best": best}
        log("class %d (%d pairs): %s -> %d lanes" % (c, len(idx), report[c]["ms"], best))
    # classes the workload leaves (nearly) empty take the choice of the same
    # size with the other hull count, else of the next smaller tuned size
    tuned = set(report)
    for c in range(NUM_CLASSES):
        if c in tuned or c % 8 == 0:
            continue
        other = (c + 8) % NUM_CLASSES
        if other in tuned:
            lanes[c] = lanes[other]
        else:
            smaller = [k for k in tuned if k // 8 == c // 8 and k % 8 < c % 8]
            if smaller:
                lanes[c] = lanes[max(smaller)]
    return lanes, report


class StepTimer:
    """Times the whole GJK + EPA step over every pair of a workload (all
    classes in one launch, as bench.py runs it); the device scene and
    buffers are built once."""

    def __init__(self, workload):
        import torch
        from . import kernels
        self.kernels = kernels
        self.dev = workload.scene.to_device()
        self.pa = torch.from_numpy(workload.pair_a).cuda()
        self.pb = torch.from_numpy(workload.pair_b).cuda()
        n = self.pa.numel()
        self.out = kernels.GjkEpaResult.empty(n, "cuda")
        self.ws = torch.empty(kernels.gjk_epa_workspace_bytes(n), dtype=torch.uint8,
                              device="cuda")

    def __call__(self, lanes, launch=None, reps=5):
        """Best of ``reps`` step times in ms."""
        import torch
        run = lambda: self.kernels.gjk_epa_batch(self.dev, self.pa, self.pb, out=self.out,
                                                 workspace=self.ws, lanes_per_pair=lanes,
                                                 launch=launch)
        run()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return 1e3 * best


def combined_step_ms(workload, lanes, reps=5, launch=None):
    """Best time of the whole GJK + EPA step over every pair of the workload."""
    return StepTimer(workload)(lanes, launch, reps)


def tune_launch(workload, lanes, sweeps=1, log=print, timer=None):
    """Launch shape minimising the whole step (coordinate descent over
    LAUNCH_CANDIDATES from LAUNCH_DEFAULT)."""
    timer = timer or StepTimer(workload)
    measure = lambda cfg: round(timer(lanes, cfg), 4)
    return descend(measure, LAUNCH_DEFAULT, LAUNCH_CANDIDATES, sweeps=sweeps,
                   log=lambda m: log("launch " + m))


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

    lanes = _native.expand_lanes(start or _native.DEFAULT_LANES)
    measure(lanes)  # warm-up
    report = {"start": {"lanes": list(lanes), "ms": round(1e3 * measure(lanes), 4)}}
    for sweep in range(sweeps):
        for c in range(NUM_CLASSES):
            if c == 8:
                continue  # two hulls always hold vertices
            times = {}
            for g in LATENCY_CANDIDATES[c]:
                trial = list(lanes)
                trial[c] = g
                times[g] = measure(trial)
            lanes[c] = pick(times)
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
    timer = StepTimer(w)
    if installed is not None and installed != lanes:
        t_new, t_old = timer(lanes), timer(installed)
        report["combined_ms"] = {"tuned": [lanes, round(t_new, 4)],
                                 "installed": [installed, round(t_old, 4)]}
        print("combined step: tuned %s %.3f ms, installed %s %.3f ms" % (lanes, t_new, installed, t_old))
        if t_old < t_new:
            lanes = installed
    launch, lreport = tune_launch(w, lanes, timer=timer)
    scene, poses, reqs = workloads.serving_workload(n_requests=20000)
    big = [r for r in reqs if r.n_pairs >= 2048][:args.serving_requests]
    slanes, sreport = tune_latency(scene, poses, big, start=lanes)
    with open(args.out, "w") as f:
        json.dump({"lanes_per_pair": lanes, "workload": "mixed_pairs seed 1, %d pairs" % args.pairs,
                   "classes": report,
                   "launch": launch, "launch_tuning": lreport,
                   "serving_lanes_per_pair": slanes,
                   "serving_workload": "%d requests of >= 2048 pairs from the serving stream "
                                       "(seed 4), one batch each on an idle GPU" % len(big),
                   "serving": sreport}, f, indent=1)
    print(json.dumps({"lanes_per_pair": lanes, "launch": launch,
                      "serving_lanes_per_pair": slanes}))


if __name__ == "__main__":
    main()
