"""Benchmark: batched GJK+EPA pair-queries/s on one MI355X.

Headline workload: 1,048,576 mixed-primitive pairs (sphere / capsule /
cylinder / box / ellipsoid / convex hull with log-uniform 8..4096 vertices),
seed 1, scene and pairs resident in HBM before the timed region.  A step is
one pass of the GJK+EPA pipeline over all pairs: GJK distance and closest
points for every pair, EPA depth / normal / contact points for the
intersecting ones.

The other synthetic configurations are measured into the same JSON line
("configs"): 4096 hull pairs, 262,144 EPA pairs, the broad phase over
1,048,576 boxes with 65,536 capsule queries plus GJK on the candidates, and
the continuous-batching server on a Poisson request stream.  Each carries its
parity against the FP64 oracle.  The oracle timed on the host cores is the
CPU baseline.
"""
import argparse
import json
import os
import sys
import time

# HIP places the process's streams round-robin on GPU_MAX_HW_QUEUES hardware
# queues (4 by default).  The serving leg keeps four batch slots in flight on
# four streams beside the default stream; with fewer queues than streams two
# slots share a queue and one batch waits behind the other.  Raised (never
# lowered) before anything initialises HIP.
try:
    _queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _queues = 4
if _queues < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "GJK+EPA pair-queries/s at ≤1e-5 rel err vs FP64; p50/p99 request latency"
MAXF = np.finfo(float).max


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
    p.add_argument("--no-parity", action="store_true", help="skip the oracle parity checks")
    p.add_argument("--no-extra", action="store_true", help="only the headline workload")
    p.add_argument("--only", type=str, default=None,
                   help="comma list of extra configs: hull,epa,broad,serving,mpr,tet,contact")
    p.add_argument("--cpu-runs", type=int, default=5)
    p.add_argument("--serving-requests", type=int, default=50000)
    p.add_argument("--classes", action="store_true",
                   help="only the per-work-class breakdown of the headline workload: each "
                        "class's pairs timed as a GJK batch of their own")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="before the workload, run one device-to-device copy of this many MiB "
                        "(a kernel of known traffic for scaling profiler byte counters)")
    return p.parse_args()


KERNEL_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                              "gjk_epa_mixed_pairs_counters.json")


def kernel_profile():
    """Per-launch counters of the headline step's kernels from the rocprofv3
    passes summarised in profiles/ (HBM bytes from FETCH_SIZE / WRITE_SIZE
    scaled by a copy kernel of known size; FP64 VALU lane operations), or
    None."""
    if not os.path.exists(KERNEL_PROFILE):
        return None
    with open(KERNEL_PROFILE) as f:
        return json.load(f)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def set_stage_timing(enable):
    from distance3d_amd import _native
    _native.check(_native.lib().d3d_stage_timing(1 if enable else 0), "d3d_stage_timing")


def stage_timing_read():
    """Mean per-call milliseconds of the GJK pipeline's stages recorded since
    stage timing was enabled (HIP events the library records on the launch
    stream): counting sort, gjk_all_kernel, epa_kernel."""
    import ctypes
    from distance3d_amd import _native
    ms = (ctypes.c_double * 3)()
    calls = ctypes.c_int64(0)
    _native.check(_native.lib().d3d_stage_timing_read(ms, ctypes.byref(calls)),
                  "d3d_stage_timing_read")
    k = max(calls.value, 1)
    return {"sort_ms": ms[0] / k, "gjk_ms": ms[1] / k, "epa_ms": ms[2] / k, "calls": calls.value}


def timed_steps(step, steps, warmup, stages=False):
    """Runs warmup + timed steps; returns (wall ms/step, event ms per step)
    and, with ``stages``, the library's per-stage times as well."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    if stages:
        set_stage_timing(True)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        ev[k][0].record(stream)
        step()
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ev_ms = np.array([a.elapsed_time(b) for a, b in ev])
    if stages:
        st = stage_timing_read()
        set_stage_timing(False)
        return wall, ev_ms, st
    return wall, ev_ms


def device_copy_bandwidth(mib=1024, reps=10):
    """Measured HBM bandwidth of a plain device-to-device copy (torch's copy
    kernel): bytes read + written per copy over its event time, median of
    ``reps``."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        b.synchronize()
        ms.append(a.elapsed_time(b))
    del src, dst
    torch.cuda.empty_cache()
    t = float(np.median(ms))
    return {"GBps": round(2.0 * (mib << 20) / (t / 1e3) / 1e9, 1), "MiB": mib, "ms": round(t, 4),
            "note": "torch device-to-device copy, read + write bytes over the event time, "
                    "median of %d" % reps}


def rel_err_stats(gpu, ref):
    sep = (ref > 1e-9) & (ref < MAXF)
    rel = np.abs(gpu[sep] - ref[sep]) / ref[sep]
    band = (ref < 1e-9) | (gpu < 1e-9)
    return {
        "max_rel_dist_err": float(rel.max(initial=0.0)),
        "p99_rel_dist_err": float(np.percentile(rel, 99)) if len(rel) else 0.0,
        "bit_identical_frac": float(np.mean(gpu == ref)),
        "flag_mismatches_outside_1e-9_band": int(np.sum(((gpu == 0.0) != (ref == 0.0)) & ~band)),
        "intersecting_frac": float(np.mean(ref == 0.0)),
    }


def cp_err_stats(cpa, cpb, ref):
    """Closest-point errors against the oracle over the pairs it does not
    clip: per pair the larger max-norm error of the two points, absolute and
    relative to max(1, |p|_inf) (gjk/_gjk_jolt.py:213-221 defines the points)."""
    ok = ref["dist"] < MAXF
    if not ok.any():
        return {"max_cp_err": 0.0, "p99_cp_err": 0.0, "max_rel_cp_err": 0.0}
    e = np.maximum(np.abs(cpa[ok] - ref["cp_a"][ok]).max(axis=1),
                   np.abs(cpb[ok] - ref["cp_b"][ok]).max(axis=1))
    scale = np.maximum(1.0, np.maximum(np.abs(ref["cp_a"][ok]).max(axis=1),
                                       np.abs(ref["cp_b"][ok]).max(axis=1)))
    return {"max_cp_err": float(e.max()), "p99_cp_err": float(np.percentile(e, 99)),
            "max_rel_cp_err": float((e / scale).max())}


EPA_STATUS = {0: "iteration_cap", 1: "converged", 2: "face_cap", 3: "degenerate_seed"}


def epa_status_counts(st):
    return {EPA_STATUS[k]: int(np.sum(st == k)) for k in EPA_STATUS}


def epa_unconverged_report(w, dev, idx, epa_out, max_iterations=64):
    """What happens on the EPA pairs that stop without converging: the same
    pairs rerun with 16x the iteration cap (how much the capped depth moves),
    and MPR's penetration depth (mpr.py:52-109) on them for comparison."""
    import torch
    from distance3d_amd import kernels
    from distance3d_amd.scene import collider_centers
    if len(idx) == 0:
        return {"pairs": 0}
    pa, pb = w.pair_a[idx], w.pair_b[idx]
    g = kernels.gjk_distance_batch(dev, pa, pb, want_simplex=True)
    long_ = kernels.epa_batch(dev, pa, pb, g.simplex, g.n_points,
                              max_iterations=16 * max_iterations)
    d_cap = epa_out[idx, 0]
    d_long = long_.depth.cpu().numpy()
    st_long = long_.status.cpu().numpy()
    centers = collider_centers(w.scene, only=np.union1d(pa, pb))
    mst, mout = kernels.mpr_batch(dev, pa, pb, centers)
    torch.cuda.synchronize()
    mst = mst.cpu().numpy()
    d_mpr = mout.cpu().numpy()[:, 0]
    hit = (mst & 1) != 0
    conv = st_long == 1
    rel = np.abs(d_cap[conv] - d_long[conv]) / np.maximum(d_long[conv], 1e-300)
    ratio = d_cap[hit] / np.maximum(d_mpr[hit], 1e-300)
    return {"pairs": int(len(idx)),
            "rerun_max_iterations": 16 * max_iterations,
            "rerun_status_counts": epa_status_counts(st_long),
            "capped_vs_rerun_max_rel_depth_diff": float(rel.max(initial=0.0)),
            "capped_vs_rerun_median_rel_depth_diff": float(np.median(rel)) if len(rel) else 0.0,
            "mpr_intersecting": int(hit.sum()),
            "capped_depth_over_mpr_depth_median": float(np.median(ratio)) if len(ratio) else None,
            "capped_depth_over_mpr_depth_min": float(ratio.min()) if len(ratio) else None,
            "capped_depth_over_mpr_depth_max": float(ratio.max()) if len(ratio) else None}


def epa_parity(gpu_out, gpu_status, ref):
    conv = ref["status"] == 1
    d, od = gpu_out[:, 0], ref["depth"]
    rel = np.abs(d[conv] - od[conv]) / np.maximum(np.abs(od[conv]), 1e-300)
    return {
        "pairs": int(len(od)),
        "status_equal_frac": float(np.mean(gpu_status == ref["status"])),
        "converged_frac": float(np.mean(conv)),
        "max_rel_depth_err": float(rel.max(initial=0.0)),
        "p99_rel_depth_err": float(np.percentile(rel, 99)) if len(rel) else 0.0,
        "max_normal_err": float(np.abs(gpu_out[conv, 1:4] - ref["normal"][conv]).max(initial=0.0)),
        "max_contact_err": float(max(
            np.abs(gpu_out[conv, 7:10] - ref["cp_a"][conv]).max(initial=0.0),
            np.abs(gpu_out[conv, 10:13] - ref["cp_b"][conv]).max(initial=0.0))),
        "bit_identical_depth_frac": float(np.mean(d == od)),
        "status_counts": epa_status_counts(gpu_status),
    }


def run_pipeline_config(w, args, lanes, steps, warmup, label, parity=True, cpu=False,
                        launch=None):
    """GJK+EPA pipeline over a workload: timing, roofline, parity, CPU."""
    import torch
    from distance3d_amd import kernels
    dev = w.scene.to_device("cuda:0")
    pa = torch.from_numpy(w.pair_a).cuda()
    pb = torch.from_numpy(w.pair_b).cuda()
    n = w.n_pairs
    out = kernels.GjkEpaResult.empty(n, "cuda")
    ws = torch.empty(kernels.gjk_epa_workspace_bytes(n), dtype=torch.uint8, device="cuda")

    def step():
        kernels.gjk_epa_batch(dev, pa, pb, out=out, workspace=ws, lanes_per_pair=lanes,
                              launch=launch)

    wall, ev_ms, stages = timed_steps(step, steps, warmup, stages=True)
    value = n / (wall / 1e3)
    # one more (untimed) pass for the per-pair GJK iteration counts
    iters = torch.empty(n, dtype=torch.int32, device="cuda")
    kernels.gjk_epa_batch(dev, pa, pb, out=out, workspace=ws, lanes_per_pair=lanes,
                          launch=launch, iterations=iters)
    iters = iters.cpu().numpy().astype(np.int64)
    dist = out.dist.cpu().numpy()
    hit = dist == 0.0
    # algorithmic bytes per pair: 12 (n_A + n_B) FP32 hull bytes, 2 x 64 pose
    # and parameters, 8 pair ids, 60 GJK outputs, +80 EPA outputs if hit
    hull_bytes = w.hull_bytes_per_pair()
    bytes_step = float(np.sum(hull_bytes) + n * (128 + 8 + 60) + 80 * hit.sum())
    achieved = bytes_step / (ev_ms.mean() / 1e3) / 1e9
    # the GJK kernel alone: hull bytes, records, pair ids and GJK outputs
    gjk_bytes = float(np.sum(hull_bytes) + n * (128 + 8 + 60))
    gjk_ms = stages["gjk_ms"]
    gjk_achieved = gjk_bytes / (gjk_ms / 1e3) / 1e9
    # every GJK iteration evaluates both supports: a full sweep of both hulls
    # would read 12 (n_A + n_B) bytes
    sweep_bytes = float(np.sum(iters * hull_bytes))
    res = {"workload": label, "pairs": n, "value": round(value, 1), "unit": "pair-queries/s",
           "ms_per_step": round(wall, 4), "event_ms_per_step": round(float(ev_ms.mean()), 4),
           "stages_ms": {k: round(v, 4) for k, v in stages.items()},
           "roofline": {"bound": "hbm", "kernel": "gjk_all_kernel",
                        "achieved": round(gjk_achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gjk_achieved / HBM_PEAK_GBS, 5),
                        "bytes_per_launch": gjk_bytes, "kernel_ms": round(gjk_ms, 4),
                        "timing": "HIP events recorded by the library on the launch stream "
                                  "around gjk_all_kernel, averaged over the timed steps",
                        "step": {"achieved": round(achieved, 2),
                                 "frac": round(achieved / HBM_PEAK_GBS, 5),
                                 "bytes_per_step": bytes_step,
                                 "ms": round(float(ev_ms.mean()), 4)}},
           "gjk_iterations": {"mean": round(float(iters.mean()), 3), "max": int(iters.max()),
                              "total": int(iters.sum())},
           "effective_sweep": {"bytes": sweep_bytes,
                               "GBps_over_step": round(sweep_bytes / (ev_ms.mean() / 1e3) / 1e9, 1),
                               "note": "sum over pairs of GJK iterations x 12 (n_A + n_B): the "
                                       "bytes full support sweeps of both hulls at every "
                                       "iteration would read, over the step time"}}
    epa_st = out.epa_status.cpu().numpy()
    res["epa_status_counts"] = epa_status_counts(epa_st)
    if parity:
        from oracle import oracle
        t0 = time.time()
        o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b, want_simplex=True)
        res["parity"] = rel_err_stats(dist, o["dist"])
        res["parity"].update(cp_err_stats(out.cp_a.cpu().numpy(), out.cp_b.cpu().numpy(), o))
        idx = np.nonzero(o["dist"] == 0.0)[0]
        if len(idx):
            e = oracle.epa(w.scene, w.pair_a[idx], w.pair_b[idx], o["simplex"][idx],
                           o["n_points"][idx])
            res["parity"]["epa"] = epa_parity(out.epa_out.cpu().numpy()[idx],
                                              out.epa_status.cpu().numpy()[idx], e)
        log("%s parity in %.1fs" % (label, time.time() - t0))
    unconv = np.nonzero((epa_st == 0) | (epa_st == 2))[0]
    if len(unconv):
        res["epa_unconverged"] = epa_unconverged_report(w, dev, unconv, out.epa_out.cpu().numpy())
    if cpu:
        res["cpu_baseline"], res["cpu_single_core"] = cpu_baseline(w, args)
    return res


def cpu_baseline(w, args):
    """The oracle (C FP64 restatement, OpenMP over pairs) on this host: all
    threads over the whole workload, 1 warm-up + median of ``--cpu-runs``; and
    one thread over the first 131,072 pairs, median of 3."""
    from oracle import oracle

    def cpu_step(pa, pb):
        o = oracle.gjk_distance(w.scene, pa, pb, want_simplex=True)
        idx = np.nonzero(o["dist"] == 0.0)[0]
        if len(idx):
            oracle.epa(w.scene, pa[idx], pb[idx], o["simplex"][idx], o["n_points"][idx])

    def timed(pa, pb, runs):
        cpu_step(pa, pb)
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            cpu_step(pa, pb)
            times.append(time.perf_counter() - t0)
        return len(pa) / float(np.median(times)), sum(times)

    cores = oracle.num_threads()
    rate, spent = timed(w.pair_a, w.pair_b, args.cpu_runs)
    full = {"value": round(rate, 1), "unit": "pair-queries/s", "cores": cores, "kind": "port",
            "sample": "all %d pairs of the same workload (GJK + EPA on intersecting pairs), "
                      "1 warm-up + median of %d runs (%.1f s); C FP64 restatement, OpenMP "
                      "dynamic schedule over pairs" % (w.n_pairs, args.cpu_runs, spent)}
    sub = min(w.n_pairs, 131072)
    oracle.set_threads(1)
    try:
        rate1, spent1 = timed(w.pair_a[:sub], w.pair_b[:sub], 3)
    finally:
        oracle.set_threads(cores)
    single = {"value": round(rate1, 1), "unit": "pair-queries/s", "cores": 1, "kind": "port",
              "sample": "first %d pairs, 1 warm-up + median of 3 runs (%.1f s)" % (sub, spent1)}
    log("cpu baseline: %.3g pairs/s on %d threads, %.3g pairs/s on 1" % (rate, cores, rate1))
    return full, single


def one_pair_per_launch(args, n_calls=2000):
    """The same library entry point called with one pair per launch and a
    synchronisation after each call, over the 4096 hull pairs of the 32-vertex
    workload (the plain per-pair GPU path)."""
    import torch
    from distance3d_amd import kernels, workloads
    w = workloads.hull_pairs()
    dev = w.scene.to_device("cuda:0")
    pa = torch.from_numpy(w.pair_a).cuda()
    pb = torch.from_numpy(w.pair_b).cuda()
    out = kernels.GjkResult(torch.empty(1, dtype=torch.float64, device="cuda"),
                            torch.empty((1, 3), dtype=torch.float64, device="cuda"),
                            torch.empty((1, 3), dtype=torch.float64, device="cuda"),
                            torch.empty(1, dtype=torch.int32, device="cuda"), None, None, None)
    ws = torch.empty(kernels.gjk_workspace_bytes(1), dtype=torch.uint8, device="cuda")
    n = w.n_pairs
    for i in range(20):
        kernels.gjk_distance_batch(dev, pa[i:i + 1], pb[i:i + 1], out=out, workspace=ws)
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(n_calls):
        k = i % n
        kernels.gjk_distance_batch(dev, pa[k:k + 1], pb[k:k + 1], out=out, workspace=ws)
        torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": round(n_calls / dt, 1), "unit": "pair-queries/s",
            "sample": "%d calls of d3d_gjk_distance_batch_ws with one pair each, synchronised "
                      "after every call, cycling over the 4096 32-vertex hull pairs" % n_calls}

    return res


def run_broad(args):
    import torch
    from distance3d_amd import kernels, workloads
    t0 = time.time()
    w = workloads.broad_phase_scene()
    nb = w.meta["n_boxes"]
    dev = w.scene.to_device("cuda:0")
    log("broad phase workload in %.1fs" % (time.time() - t0))
    ids = torch.arange(w.scene.n_shapes, dtype=torch.int32, device="cuda")
    aabbs = kernels.aabb_batch(dev, ids)
    buf = torch.empty(int(kernels._native.load_library().d3d_bvh_bytes(nb)), dtype=torch.uint8,
                      device="cuda")
    cap = 1 << 21
    pairs = torch.empty((cap, 2), dtype=torch.int32, device="cuda")
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    lib = kernels.lib()
    P = kernels.ptr
    box_ptr = aabbs.data_ptr()
    q_ptr = aabbs[nb:].data_ptr()

    def build():
        kernels.check(lib.d3d_aabb_batch(dev.c, P(ids), w.scene.n_shapes, P(aabbs),
                                         kernels.stream_ptr()), "aabb")
        kernels.check(lib.d3d_bvh_build(box_ptr, nb, P(buf), buf.numel(), kernels.stream_ptr()),
                      "build")

    def query():
        kernels.check(lib.d3d_bvh_query(P(buf), nb, box_ptr, q_ptr, w.scene.n_shapes - nb,
                                        P(pairs), cap, P(counter), kernels.stream_ptr()), "query")

    def refit():
        kernels.check(lib.d3d_aabb_batch(dev.c, P(ids), w.scene.n_shapes, P(aabbs),
                                         kernels.stream_ptr()), "aabb")
        kernels.check(lib.d3d_bvh_refit(box_ptr, nb, P(buf), buf.numel(), kernels.stream_ptr()),
                      "refit")

    b_wall, b_ev = timed_steps(build, 10, 2)
    r_wall, r_ev = timed_steps(refit, 10, 2)
    q_wall, q_ev = timed_steps(query, 10, 2)
    k = int(counter.item())
    cand = pairs[:k].cpu().numpy()
    ca = (nb + cand[:, 0]).astype(np.int32)
    cb = cand[:, 1].astype(np.int32)
    wn = workloads.Workload("broad_phase_candidates", w.scene, ca, cb)
    narrow = run_pipeline_config(wn, args, None, 10, 2, "broad-phase candidates (capsule x box)",
                                 parity=not args.no_parity)
    # leaf AABBs 48 B + queries 48 B + 8 B per emitted pair
    q_bytes = nb * 48 + (w.scene.n_shapes - nb) * 48 + 8 * k
    res = {"workload": "broad_phase (1,048,576 boxes, 65,536 capsule queries)",
           "boxes": nb, "queries": w.scene.n_shapes - nb, "candidate_pairs": k,
           "build_ms": round(float(b_ev.mean()), 4),
           "refit_ms": round(float(r_ev.mean()), 4),
           "query_ms": round(float(q_ev.mean()), 4),
           "query_roofline": {"bound": "hbm", "achieved": round(q_bytes / (q_ev.mean() / 1e3) / 1e9, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s"},
           "narrow_phase": narrow}
    if not args.no_parity:
        from oracle import oracle
        t0 = time.time()
        host = aabbs.cpu().numpy()
        o, cnt = oracle.overlap_pairs_sorted(host[nb:], host[:nb], cap=cap)
        g = cand[np.lexsort((cand[:, 1], cand[:, 0]))].astype(np.int64)
        o = o[np.lexsort((o[:, 1], o[:, 0]))]
        res["parity"] = {"pair_set_equal": bool(cnt == k and np.array_equal(g, o)),
                         "oracle_pairs": int(cnt),
                         "aabbs_bitwise_equal": bool(np.array_equal(
                             host, oracle.aabbs(w.scene, np.arange(w.scene.n_shapes))))}
        log("broad parity in %.1fs" % (time.time() - t0))
    return res


def run_serving(args, lanes):
    from distance3d_amd import autotune, serving, workloads
    t0 = time.time()
    scene, poses, reqs = workloads.serving_workload(n_requests=args.serving_requests)
    log("serving workload in %.1fs: %d requests, %d pairs"
        % (time.time() - t0, len(reqs), sum(r.n_pairs for r in reqs)))
    dev = scene.to_device("cuda:0")
    srv = serving.CollisionServer(dev, poses, n_slots=4, slot_pairs=16384,
                                  lanes_per_pair=autotune.tuned_lanes(kind="latency"))
    log("server ready (graphs captured)")
    # warm-up: the first 2000 requests of the stream, served and discarded
    srv.serve(reqs[:2000], realtime=True)
    res, stats = srv.serve(reqs, realtime=True)
    stats["warmup_requests"] = 2000
    stats["slots"] = 4
    stats["lanes_per_pair"] = srv.lanes_per_pair
    stats["workload"] = ("Poisson stream, %d requests at 20k req/s, log-uniform 1..4096 "
                         "pose-updated pairs, 512-collider resident scene" % len(reqs))
    stats["slot_bytes"] = serving.CollisionServer.slot_bytes(16384)
    if not args.no_parity:
        from oracle import oracle
        worst = 0.0
        flags = 0
        checked = 0
        for r in reqs[:: max(1, len(reqs) // 200)]:
            host = workloads.posed_instances(scene, poses, r.pair_a, r.pair_b, r.pose_a, r.pose_b)
            k = r.n_pairs
            g = oracle.gjk_distance(host, np.arange(0, 2 * k, 2), np.arange(1, 2 * k, 2))
            d = res[r.rid].dist
            flags += int(np.sum((d == 0.0) != (g["dist"] == 0.0)))
            sep = g["dist"] > 1e-9
            if sep.any():
                worst = max(worst, float(np.max(np.abs(d[sep] - g["dist"][sep]) / g["dist"][sep])))
            checked += k
        stats["parity"] = {"pairs_checked": checked, "max_rel_dist_err": worst,
                           "flag_mismatches": flags}
    return stats


def run_mpr(w, args, steps=5, warmup=1):
    """MPR penetration (mpr.py:52-109) over the headline's pairs; parity of a
    sample against the oracle."""
    import torch
    from distance3d_amd import kernels
    from distance3d_amd.scene import collider_centers
    t0 = time.time()
    centers = collider_centers(w.scene)
    dev = w.scene.to_device("cuda:0")
    log("mpr centres and scene in %.1fs" % (time.time() - t0))
    n = w.n_pairs
    pa = torch.from_numpy(w.pair_a).cuda()
    pb = torch.from_numpy(w.pair_b).cuda()
    c = torch.from_numpy(centers).cuda()
    st = torch.empty(n, dtype=torch.int32, device="cuda")
    out = torch.empty((n, 7), dtype=torch.float64, device="cuda")
    lib = kernels.lib()
    P = kernels.ptr

    def step():
        kernels.check(lib.d3d_mpr_batch(dev.c, P(pa), P(pb), P(c), n, 0.0001, 100, 1, P(st), P(out),
                                        kernels.stream_ptr()), "mpr")

    wall, ev = timed_steps(step, steps, warmup)
    ms = float(ev.mean())
    res = {"workload": "mpr on mixed_pairs (1,048,576 mixed-primitive pairs, MPR penetration)",
           "pairs": n, "value": round(n / (ms / 1e3), 1), "unit": "pair-queries/s",
           "ms_per_step": round(ms, 4), "intersecting_frac": float((st == 1).float().mean())}
    if not args.no_parity:
        from oracle import oracle
        idx = np.random.RandomState(7).choice(n, 20000, replace=False)
        est, eout = oracle.mpr(w.scene, w.pair_a[idx], w.pair_b[idx], centers)
        gst = st.cpu().numpy()[idx]
        gout = out.cpu().numpy()[idx]
        same = np.all((gout == eout) | (np.isnan(gout) & np.isnan(eout)), axis=1)
        res["parity"] = {"pairs_checked": len(idx), "status_equal_frac": float(np.mean(gst == est)),
                         "bit_identical_frac": float(np.mean(same))}
    log("mpr: %.3f ms/step" % ms)
    return res


def run_tet(args, steps=5, warmup=1):
    """Hydroelastic tetrahedron-pair contact (contact plane and polygon) for
    1,048,576 random pairs between two bodies of 100,000 tetrahedra each."""
    import torch
    from distance3d_amd import hydroelastic_contact as hc, kernels
    rs = np.random.RandomState(10)
    t0 = time.time()
    n, m = 100000, 1 << 20
    T = [rs.uniform(-0.6, 0.6, (n, 1, 3)) + rs.randn(n, 4, 3) * 0.5 for _ in range(2)]
    E = [rs.uniform(0.0, 1.0, (n, 4)) for _ in range(2)]
    X = [hc.barycentric_transforms(t) for t in T]
    pairs = np.stack([rs.randint(0, n, m), rs.randint(0, n, m)], 1).astype(np.int32)
    log("tet workload in %.1fs" % (time.time() - t0))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    T1, T2, e1, e2, X1, X2, pr = map(d, (T[0], T[1], E[0], E[1], X[0], X[1], pairs))
    st = torch.empty(m, dtype=torch.int32, device="cuda")
    planes = torch.empty((m, 4), dtype=torch.float64, device="cuda")
    nv = torch.empty(m, dtype=torch.int32, device="cuda")
    polys = torch.empty((m, kernels.TET_MAX_POLY, 3), dtype=torch.float64, device="cuda")
    ws = torch.empty(int(kernels.lib().d3d_tet_pairs_workspace_bytes(m)), dtype=torch.uint8,
                     device="cuda")
    lib = kernels.lib()
    P = kernels.ptr

    def step():
        kernels.check(lib.d3d_tet_pairs_batch(P(T1), P(e1), P(X1), P(T2), P(e2), P(X2), P(pr), m,
                                              1.0, 1.0, P(st), P(planes), P(nv), P(polys), P(ws),
                                              ws.numel(), kernels.stream_ptr()), "tet")

    wall, ev = timed_steps(step, steps, warmup)
    ms = float(ev.mean())
    res = {"workload": "tetrahedron pairs (1,048,576 random pairs, two bodies of 100,000 "
                       "tetrahedra): contact plane and polygon",
           "pairs": m, "value": round(m / (ms / 1e3), 1), "unit": "pair-queries/s",
           "ms_per_step": round(ms, 4), "intersecting_frac": float((st & 1).float().mean())}
    if not args.no_parity:
        from oracle import oracle
        idx = np.random.RandomState(11).choice(m, 20000, replace=False)
        want = oracle.tet_pairs(T[0], E[0], X[0], T[1], E[1], X[1], pairs[idx])
        got = [t.cpu().numpy()[idx] for t in (st, planes, nv, polys)]
        res["parity"] = {"pairs_checked": len(idx),
                         "status_equal_frac": float(np.mean(got[0] == want[0])),
                         "plane_bit_identical_frac": float(np.mean(np.all(got[1] == want[1], 1))),
                         "polygon_bit_identical_frac": float(np.mean(np.all(
                             (got[3] == want[3]) | (np.isnan(got[3]) & np.isnan(want[3])),
                             axis=(1, 2))))}
    log("tet: %.3f ms/step" % ms)
    return res


def run_contact(args, steps=10, warmup=2):
    """Hydroelastic contact surface of two overlapping icosphere bodies
    (order 4: 5,120 tetrahedra each, centres 0.14 apart at radius 0.15):
    tetrahedron boxes and barycentric transforms, brute-force overlap of the
    boxes (26 M box tests, 5.1 M candidate pairs, row-major as the
    reference lists them), status pass, compaction of the
    intersecting pairs, polygon and force pass.  Meshes resident in HBM; the
    step includes its two host reads (pair count, hit count)."""
    import torch
    from distance3d_amd import hydroelastic_contact as hc, kernels
    rb1 = hc.RigidBody.make_sphere(np.array([0.0, 0.0, 0.01]), 0.15, 4)
    rb2 = hc.RigidBody.make_sphere(np.array([0.0, 0.0, 0.15]), 0.15, 4)
    rb1.express_in(rb2.body2origin_)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    T1, T2 = d(rb1.tetrahedra_points), d(rb2.tetrahedra_points)
    e1, e2 = d(rb1.tetrahedra_potentials), d(rb2.tetrahedra_potentials)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    out = {}

    def step():
        ab1, X1, _ = kernels.tet_mesh_prep(T1)
        ab2, X2, _ = kernels.tet_mesh_prep(T2)
        pairs = kernels.overlap_dense(ab1, ab2, cap=6000000, counter=counter)
        st = kernels.tet_pairs_batch(T1, e1, X1, T2, e2, X2, pairs, 1.0, 1.0, geometry=False,
                                     checked=False)[0]
        rows = torch.nonzero(st & 0x101).squeeze(1)
        sel = st[rows]
        if bool((sel & 0x100).any()):
            raise AssertionError("half-plane overflow")
        hp = pairs[rows[(sel & 1) != 0]]
        st2, planes, nv, polys = kernels.tet_pairs_batch(T1, e1, X1, T2, e2, X2, hp, 1.0, 1.0,
                                                         checked=False)
        F = kernels.tet_contact_forces(T1, e1, hp, st2, planes, nv, polys, 1.0, checked=False)
        out.update(pairs=pairs, hp=hp, planes=planes, nv=nv, polys=polys, F=F)

    wall, ev = timed_steps(step, steps, warmup)
    n_pairs, n_hits = int(out["pairs"].shape[0]), int(out["hp"].shape[0])
    res = {"workload": "hydroelastic contact surface, two order-4 icosphere bodies (5,120 "
                       "tetrahedra each, deep overlap): prep, broad phase, pair contact, forces",
           "tetrahedra": [int(T1.shape[0]), int(T2.shape[0])], "candidate_pairs": n_pairs,
           "intersecting_pairs": n_hits, "ms_per_step": round(wall, 4),
           "value": round(n_pairs / (wall / 1e3), 1), "unit": "candidate tetrahedron pairs/s"}
    if not args.no_parity:
        from oracle import oracle
        t0 = time.perf_counter()
        P1, P2 = rb1.tetrahedra_points, rb2.tetrahedra_points
        a1, Y1, _ = oracle.tet_mesh_prep(P1)
        a2, Y2, _ = oracle.tet_mesh_prep(P2)
        cap = n_pairs + 1
        wp = np.empty((cap, 2), dtype=np.int64)
        k = oracle.lib().orc_all_aabbs_overlap(oracle._p(a1), len(a1), oracle._p(a2), len(a2),
                                               oracle._p(wp), cap)
        wp = wp[:k]
        st, pl, nvw, po = oracle.tet_pairs(P1, rb1.tetrahedra_potentials, Y1, P2,
                                           rb2.tetrahedra_potentials, Y2, wp)
        Fw = oracle.tet_contact_forces(P1, rb1.tetrahedra_potentials, wp, st, pl, nvw, po)
        cpu_s = time.perf_counter() - t0
        hit = np.nonzero(st & 1)[0]
        same = (k == n_pairs and np.array_equal(out["pairs"].cpu().numpy(), wp)
                and np.array_equal(out["hp"].cpu().numpy(), wp[hit])
                and np.array_equal(out["planes"].cpu().numpy(), pl[hit])
                and np.array_equal(out["nv"].cpu().numpy(), nvw[hit])
                and np.array_equal(out["polys"].cpu().numpy(), po[hit], equal_nan=True)
                and np.array_equal(out["F"].cpu().numpy(), Fw[hit]))
        res["parity"] = {"pairs_hits_planes_polygons_forces_bit_identical": bool(same)}
        res["cpu_oracle"] = {"seconds": round(cpu_s, 3), "threads": oracle.num_threads(),
                             "note": "brute-force box overlaps on one core, pair contact and "
                                     "forces on all host threads"}
    log("contact: %.3f ms/step, %d candidate pairs, %d hits" % (wall, n_pairs, n_hits))
    return res


def class_breakdown(w, lanes, reps=5, launch=None):
    """Each work class of the workload (pairs sorted by their total hull
    vertex count and whether both colliders are hulls, 16 classes) as a GJK
    batch of its own: event time, pairs,
    GJK iterations, and the algorithmic bytes 12 (n_A + n_B) + 196 per pair
    over that time.  Each class is one set of dispatches, so a kernel trace
    of this mode gives per-class kernel times and counters."""
    import torch
    from distance3d_amd import autotune, kernels
    dev = w.scene.to_device("cuda:0")
    cls = autotune.pair_classes(w.scene, w.pair_a, w.pair_b)
    hb = w.hull_bytes_per_pair()
    rows = {}
    for c in range(16):
        idx = np.nonzero(cls == c)[0]
        if len(idx) == 0:
            continue
        pa = torch.from_numpy(w.pair_a[idx]).cuda()
        pb = torch.from_numpy(w.pair_b[idx]).cuda()
        n = len(idx)
        out = kernels.GjkResult(torch.empty(n, dtype=torch.float64, device="cuda"),
                                torch.empty((n, 3), dtype=torch.float64, device="cuda"),
                                torch.empty((n, 3), dtype=torch.float64, device="cuda"),
                                torch.empty(n, dtype=torch.int32, device="cuda"),
                                torch.empty(n, dtype=torch.int32, device="cuda"), None, None)
        ws = torch.empty(kernels.gjk_workspace_bytes(n), dtype=torch.uint8, device="cuda")

        def step():
            kernels.gjk_distance_batch(dev, pa, pb, lanes_per_pair=lanes, out=out, workspace=ws,
                                       launch=launch)

        log("class %d: %d pairs" % (c, n))
        wall, ev = timed_steps(step, reps, 1)
        it = out.iterations.cpu().numpy().astype(np.int64)
        b = float(np.sum(hb[idx]) + 196 * n)
        ms = float(ev.mean())
        rows[str(c)] = {"pairs": n, "ms": round(ms, 4), "lanes": (lanes or list(
            _native_default_lanes()))[c], "iterations_mean": round(float(it.mean()), 3),
            "hull_vertices_mean": round(float(hb[idx].mean()) / 12.0, 1),
            "algorithmic_GBps": round(b / (ms / 1e3) / 1e9, 1),
            "sweep_GBps": round(float(np.sum(it * hb[idx])) / (ms / 1e3) / 1e9, 1)}
    return rows


def _native_default_lanes():
    from distance3d_amd import _native
    return _native.DEFAULT_LANES


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the collision path shards by pair range "
              "with no exchange, so a multi-GPU figure is N times this one and is not "
              "measured here (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401
    from distance3d_amd import _native, workloads

    from distance3d_amd import autotune
    lanes = [int(x) for x in args.lanes.split(",")] if args.lanes else autotune.tuned_lanes()
    launch = autotune.tuned_launch()
    if args.copy_calibration_mib > 0:
        src = torch.ones(args.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: %d MiB read + %d MiB written" % (args.copy_calibration_mib,
                                                               args.copy_calibration_mib))
        del src, dst
    copy_bw = device_copy_bandwidth()
    log("device copy: %.1f GB/s" % copy_bw["GBps"])
    t0 = time.time()
    w = workloads.mixed_pairs(seed=args.seed, n_pairs=args.pairs)
    log("workload generated in %.1fs: %d pairs, %d hull vertices"
        % (time.time() - t0, w.n_pairs, w.scene.n_vertices))
    if args.classes:
        print(json.dumps({"class_breakdown": class_breakdown(w, lanes, launch=launch),
                          "workload": "mixed_pairs seed %d, %d pairs" % (args.seed, w.n_pairs)}),
              flush=True)
        return
    head = run_pipeline_config(w, args, lanes, args.steps, args.warmup, "mixed_pairs",
                               parity=not args.no_parity, cpu=not args.no_cpu, launch=launch)
    log("headline: %.3f ms/step" % head["ms_per_step"])
    result = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "pair-queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (simplex, EPA) + f32 hull sweeps",
        "data": "synthetic",
        "config": {"workload": "mixed_pairs: 1,048,576 mixed-primitive pairs, GJK + EPA on "
                               "intersecting pairs",
                   "pairs": w.n_pairs, "seed": args.seed, "hull_vertices": "log-uniform 8..4096",
                   "lanes_per_pair": lanes or list(_native.DEFAULT_LANES),
                   "launch": launch or dict(autotune.LAUNCH_DEFAULT)},
        "roofline": dict(head["roofline"], device_copy=copy_bw,
                         frac_of_copy=round(head["roofline"]["achieved"] / copy_bw["GBps"], 5)),
    }
    prof = kernel_profile()
    stages = head["stages_ms"]
    rl = result["roofline"]
    rl["traffic"] = None
    if prof is not None:
        g = prof["kernels"]["gjk_all_kernel"]
        rl["traffic"] = g["fetch_bytes"] + g["write_bytes"]
        rl["traffic_source"] = ("profiles/gjk_epa_mixed_pairs_counters.md: rocprofv3 FETCH_SIZE and "
                                "WRITE_SIZE passes of this workload, gjk_all_kernel per launch, "
                                "scaled by a copy kernel of known size")
        rl["traffic_over_algorithmic"] = round(rl["traffic"] / rl["bytes_per_launch"], 3)
        fp = {}
        for name, ms in (("gjk_all_kernel", stages["gjk_ms"]), ("epa_kernel", stages["epa_ms"])):
            ops = prof["kernels"][name]["fp64_lane_ops"]
            tf = ops / (ms / 1e3) / 1e12
            fp[name] = {"lane_ops_per_launch": ops, "achieved": round(tf, 3), "peak": 78.6,
                        "unit": "TFLOP/s", "frac": round(tf / 78.6, 5)}
        result["fp64_valu"] = dict(fp, note=(
            "FP64 VALU operations per launch from rocprofv3 (SQ_INSTS_VALU_ADD/MUL/TRANS_F64 + "
            "2 x FMA_F64, x 64 lanes: issued lanes, an upper bound on useful FLOPs) over the "
            "kernel time measured in this run"))
    for k in ("stages_ms", "gjk_iterations", "effective_sweep", "epa_status_counts",
              "epa_unconverged"):
        if k in head:
            result[k] = head[k]
    if "parity" in head:
        result["parity"] = head["parity"]
    if "cpu_baseline" in head:
        result["cpu_baseline"] = head["cpu_baseline"]
        result["cpu_single_core"] = head["cpu_single_core"]
        result["speedup_vs_cpu"] = round(head["value"] / head["cpu_baseline"]["value"], 2)
        result["speedup_vs_cpu_single_core"] = round(
            head["value"] / head["cpu_single_core"]["value"], 2)
    if not args.no_extra:
        opl = one_pair_per_launch(args)
        result["one_pair_per_launch"] = opl
        result["speedup_vs_one_pair_per_launch"] = round(head["value"] / opl["value"], 1)
        log("one pair per launch: %.3g pairs/s" % opl["value"])
    if not args.no_extra:
        only = (set(args.only.split(",")) if args.only
                else {"hull", "epa", "broad", "serving", "mpr", "tet", "contact"})
        cfgs = {}
        if "hull" in only:
            cfgs["hull_pairs"] = run_pipeline_config(workloads.hull_pairs(), args, lanes, 20, 3,
                                                     "hull_pairs (4096 x 32-vertex hulls)",
                                                     parity=not args.no_parity, launch=launch)
        if "epa" in only:
            t0 = time.time()
            we = workloads.epa_pairs()
            log("epa workload in %.1fs" % (time.time() - t0))
            cfgs["epa_pairs"] = run_pipeline_config(we, args, lanes, 5, 1,
                                                    "epa_pairs (262,144 overlapping hull pairs)",
                                                    parity=not args.no_parity, launch=launch)
        if "broad" in only:
            cfgs["broad_phase"] = run_broad(args)
        if "serving" in only:
            cfgs["serving"] = run_serving(args, lanes)
        if "mpr" in only:
            cfgs["mpr"] = run_mpr(w, args)
        if "tet" in only:
            cfgs["tet_pairs"] = run_tet(args)
        if "contact" in only:
            cfgs["contact_surface"] = run_contact(args)
        result["configs"] = cfgs
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
