"""The synthetic configurations at their stated sizes and seeds, on the GPU,
against the FP64 oracle.

* mixed pairs: 1,048,576 mixed-primitive pairs (seed 1), GJK + EPA pipeline;
* EPA pairs: 262,144 overlapping hull pairs of 64-256 vertices (seed 2);
* broad phase: 1,048,576 oriented boxes and 65,536 query capsules (seed 3),
  exact pair set, then GJK on the candidates;
* serving: a real-time Poisson stream over the 512-collider scene (seed 4),
  reduced to 2,000 requests, each result checked alone against in the mix
  on a sample and against the oracle.

Tolerances, as in the per-kernel tests: distance within 1e-5 relative
(absolute 1e-12 near zero), closest points within 1e-5 * max(1, |p|),
intersection flags identical outside the 1e-9 contact band, EPA status equal
and depth within 1e-6 * max(1, depth), contact points within 1e-5, the
broad-phase pair set identical.
"""
import numpy as np
import pytest

from distance3d_amd import _native, kernels, serving, workloads
from oracle import oracle

pytestmark = pytest.mark.gpu

REL = 1e-5
BAND = 1e-9


def _pipeline(w):
    import torch
    dev = w.scene.to_device("cuda")
    pa = torch.from_numpy(w.pair_a).cuda()
    pb = torch.from_numpy(w.pair_b).cuda()
    out = kernels.gjk_epa_batch(dev, pa, pb)
    torch.cuda.synchronize()
    return {k: getattr(out, k).cpu().numpy()
            for k in ("dist", "cp_a", "cp_b", "status", "epa_out", "epa_status")}


def _check_pipeline(w, r):
    o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b, want_simplex=True)
    dist, od = r["dist"], o["dist"]
    band = (od < BAND) | (dist < BAND)
    assert not np.any(((dist == 0.0) != (od == 0.0)) & ~band)
    sep = (od >= BAND) & (od < _native.MAX_FLOAT)
    assert np.all(np.abs(dist[sep] - od[sep]) <= REL * od[sep])
    clipped = od == _native.MAX_FLOAT
    assert np.all(dist[clipped] == _native.MAX_FLOAT)
    ok = ~clipped
    scale = np.maximum(1.0, np.abs(o["cp_a"][ok]).max(axis=1))
    assert np.all(np.abs(r["cp_a"][ok] - o["cp_a"][ok]).max(axis=1) <= REL * scale)
    assert np.all(np.abs(r["cp_b"][ok] - o["cp_b"][ok]).max(axis=1) <= REL * scale)
    assert np.array_equal(r["status"], o["status"])
    hit = np.nonzero(od == 0.0)[0]
    assert np.all(r["epa_status"][od > 0.0] == -1)
    e = oracle.epa(w.scene, w.pair_a[hit], w.pair_b[hit], o["simplex"][hit], o["n_points"][hit])
    g = r["epa_out"][hit]
    assert np.array_equal(r["epa_status"][hit], e["status"])
    assert np.all(np.abs(g[:, 0] - e["depth"]) <= 1e-6 * np.maximum(1.0, np.abs(e["depth"])))
    conv = e["status"] == 1
    assert np.all(np.abs(g[conv, 1:4] - e["normal"][conv]) <= 1e-6)
    assert np.all(np.abs(g[conv, 7:10] - e["cp_a"][conv]) <= 1e-5)
    assert np.all(np.abs(g[conv, 10:13] - e["cp_b"][conv]) <= 1e-5)
    return o, e, hit


def test_mixed_pairs_full_size(gpu):
    w = workloads.mixed_pairs(seed=1, n_pairs=1 << 20)
    print("mixed pairs: %d pairs, %d hull vertices" % (w.n_pairs, w.scene.n_vertices), flush=True)
    o, e, hit = _check_pipeline(w, _pipeline(w))
    assert 0.02 < len(hit) / w.n_pairs < 0.05


def test_epa_pairs_full_size(gpu):
    w = workloads.epa_pairs(seed=2, n_pairs=262144)
    print("epa pairs: %d pairs" % w.n_pairs, flush=True)
    r = _pipeline(w)
    o, e, hit = _check_pipeline(w, r)
    # the generator overlaps the two hulls' extents along one axis by U(0.01,
    # 0.5); about 90% of the pairs then intersect
    assert len(hit) > 0.85 * w.n_pairs
    # a converged depth never exceeds the overlap the generator built in
    conv = e["status"] == 1
    assert np.all(r["epa_out"][hit][conv, 0] <= w.meta["delta"][hit][conv] + 1e-9)


def test_broad_phase_full_size(gpu):
    import torch
    w = workloads.broad_phase_scene(seed=3)
    nb = w.meta["n_boxes"]
    assert nb == 1 << 20 and w.scene.n_shapes - nb == 65536
    dev = w.scene.to_device("cuda")
    ids = np.arange(w.scene.n_shapes, dtype=np.int32)
    aabbs = kernels.aabb_batch(dev, ids)
    host = aabbs.cpu().numpy()
    assert np.array_equal(host, oracle.aabbs(w.scene, ids))
    bvh = kernels.bvh_build(aabbs[:nb])
    g = kernels.bvh_query(bvh, aabbs[nb:]).cpu().numpy().astype(np.int64)
    o, cnt = oracle.overlap_pairs_sorted(host[nb:], host[:nb], cap=1 << 21)
    o = o[np.lexsort((o[:, 1], o[:, 0]))]
    print("broad phase: %d candidate pairs" % cnt, flush=True)
    assert cnt == len(g) and np.array_equal(g, o)
    # narrow phase on the candidates: capsule (query) against box (leaf)
    wn = workloads.Workload("candidates", w.scene, (nb + g[:, 0]).astype(np.int32),
                            g[:, 1].astype(np.int32))
    _check_pipeline(wn, _pipeline(wn))
    del bvh
    torch.cuda.empty_cache()


def test_serving_realtime_stream_reduced(gpu):
    scene, poses, reqs = workloads.serving_workload(seed=4, n_colliders=512, n_requests=2000)
    dev = scene.to_device("cuda")
    srv = serving.CollisionServer(dev, poses, n_slots=4, slot_pairs=16384)
    mix, stats = srv.serve(reqs, realtime=True)
    assert stats["requests"] == len(reqs) and sorted(mix) == [r.rid for r in reqs]
    sample = reqs[::50]
    for r in sample[:20]:
        alone, _ = srv.serve([serving.Request(r.rid, 0.0, r.pair_a, r.pair_b, r.pose_a,
                                              r.pose_b)], realtime=False)
        assert alone[r.rid].same_as(mix[r.rid])
    for r in sample:
        host = workloads.posed_instances(scene, poses, r.pair_a, r.pair_b, r.pose_a, r.pose_b)
        k = r.n_pairs
        a, b = np.arange(0, 2 * k, 2), np.arange(1, 2 * k, 2)
        g = oracle.gjk_distance(host, a, b, want_simplex=True)
        out = mix[r.rid]
        band = (g["dist"] < BAND) | (out.dist < BAND)
        assert not np.any(((out.dist == 0.0) != (g["dist"] == 0.0)) & ~band)
        sep = g["dist"] >= BAND
        assert np.all(np.abs(out.dist[sep] - g["dist"][sep]) <= REL * g["dist"][sep])
        hit = np.nonzero(g["dist"] == 0.0)[0]
        if len(hit):
            e = oracle.epa(host, a[hit], b[hit], g["simplex"][hit], g["n_points"][hit])
            assert np.array_equal(out.epa_status[hit], e["status"])
            assert np.all(np.abs(out.depth[hit] - e["depth"]) <= 1e-6 * np.maximum(1, e["depth"]))
