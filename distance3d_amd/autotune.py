"""Launch tuning for the batched GJK: lanes per pair for each work class.

The GJK pipeline sorts pairs into 8 classes by their total hull vertex count
(0, <=32, <=128, <=512, <=1024, <=2048, <=4096, >4096).  For each class this
tuner times the GJK kernels on that class's own pairs (taken from a sample
workload) with every candidate lane count and keeps the fastest; the table is
written to ``tuning.json`` beside this file and picked up by
:func:`tuned_lanes`.

Usage: ``python -m distance3d_amd.autotune [--pairs N]`` on the GPU.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _native

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
CLASS_EDGES = (0, 32, 128, 512, 1024, 2048, 4096)
CANDIDATES = {0: (1, 2, 4), 1: (1, 2, 4), 2: (1, 2, 4, 8), 3: (2, 4, 8, 16), 4: (2, 4, 8, 16),
              5: (4, 8, 16, 32), 6: (4, 8, 16, 32, 64), 7: (8, 16, 32, 64)}


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


def tuned_lanes(path=TUNING_PATH):
    """The tuned table, or None when no tuning file exists."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        lanes = json.load(f)["lanes_per_pair"]
    return [int(x) for x in lanes]


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


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--pairs", type=int, default=1 << 20)
    p.add_argument("--out", default=TUNING_PATH)
    args = p.parse_args()
    from . import workloads
    w = workloads.mixed_pairs(seed=1, n_pairs=args.pairs)
    lanes, report = tune(w)
    with open(args.out, "w") as f:
        json.dump({"lanes_per_pair": lanes, "workload": "mixed_pairs seed 1, %d pairs" % args.pairs,
                   "classes": report}, f, indent=1)
    print(json.dumps({"lanes_per_pair": lanes}))


if __name__ == "__main__":
    main()
