"""GPU parity of the GJK / support / AABB kernels against the FP64 oracle.

Tolerances: distances within 1e-5 relative (absolute 1e-12 near zero),
closest points within 1e-5 * scale, intersection flags identical outside a
1e-9 contact band; support points and AABBs identical to the last bit (both
sides evaluate the same FP64 expressions without contraction and the hull
argmax is exact).  The kernels are reached through the C ABI only.
"""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal
from pytest import approx

from distance3d_amd import colliders, gjk, kernels, workloads
from distance3d_amd import _native
from distance3d_amd.scene import SceneBuilder
from oracle import oracle
from tests import geometry_brute as gb

pytestmark = pytest.mark.gpu

REL = 1e-5
BAND = 1e-9


def _np(t):
    return t.cpu().numpy()


def _compare_gjk(host, pa, pb, **kw):
    dev = host.to_device()
    r = kernels.gjk_distance_batch(dev, pa, pb, **kw)
    o = oracle.gjk_distance(host, pa, pb)
    dist, cpa, cpb, st = _np(r.dist), _np(r.cp_a), _np(r.cp_b), _np(r.status)
    od = o["dist"]
    # intersection flags exact outside the contact band
    band = (od < BAND) | (dist < BAND)
    flag_mismatch = ((dist == 0.0) != (od == 0.0)) & ~band
    assert not np.any(flag_mismatch), np.nonzero(flag_mismatch)[0][:10]
    sep = (od >= BAND) & (od < _native.MAX_FLOAT)
    rel = np.abs(dist[sep] - od[sep]) / od[sep]
    assert rel.max(initial=0.0) <= REL, rel.max()
    clipped = od == _native.MAX_FLOAT
    assert np.all(dist[clipped] == _native.MAX_FLOAT)
    scale = np.maximum(1.0, np.abs(o["cp_a"]).max(axis=1))
    ok = ~clipped
    assert np.all(np.abs(cpa[ok] - o["cp_a"][ok]).max(axis=1) <= REL * scale[ok])
    assert np.all(np.abs(cpb[ok] - o["cp_b"][ok]).max(axis=1) <= REL * scale[ok])
    return r, o, rel


def test_hull_pairs_full_parity(gpu):
    w = workloads.hull_pairs()
    r, o, rel = _compare_gjk(w.scene, w.pair_a, w.pair_b)
    # same FP64 arithmetic and exact supports: the iteration paths coincide
    assert np.array_equal(_np(r.iterations), o["iterations"])
    assert np.mean(_np(r.dist) == o["dist"]) > 0.99


def test_mixed_pairs_parity(gpu):
    w = workloads.mixed_pairs(seed=5, n_pairs=20000)
    r, o, rel = _compare_gjk(w.scene, w.pair_a, w.pair_b)
    assert np.mean(_np(r.dist) == o["dist"]) > 0.99


def test_mixed_pairs_all_collider_types_with_margin(gpu):
    rs = np.random.RandomState(3)
    b = SceneBuilder()
    ids = []
    for k in range(400):
        T = gb.random_pose(rs, 1.2)
        m = 0.05 if k % 5 == 0 else 0.0
        kind = k % 9
        if kind == 0:
            ids.append(b.add_sphere(T[:3, 3], rs.uniform(0.1, 1), margin=m))
        elif kind == 1:
            ids.append(b.add_capsule(T, rs.uniform(0.05, 0.5), rs.uniform(0.2, 2), margin=m))
        elif kind == 2:
            ids.append(b.add_cylinder(T, rs.uniform(0.05, 0.5), rs.uniform(0.2, 2), margin=m))
        elif kind == 3:
            ids.append(b.add_box(T, rs.uniform(0.1, 2, 3), margin=m))
        elif kind == 4:
            ids.append(b.add_ellipsoid(T, rs.uniform(0.1, 1, 3), margin=m))
        elif kind == 5:
            ids.append(b.add_cone(T, rs.uniform(0.1, 0.6), rs.uniform(0.2, 1.5), margin=m))
        elif kind == 6:
            ids.append(b.add_disk(T[:3, 3], rs.uniform(0.1, 1), T[:3, 2], margin=m))
        elif kind == 7:
            ids.append(b.add_ellipse(T[:3, 3], T[:3, :2].T, rs.uniform(0.1, 1, 2), margin=m))
        else:
            n = int(rs.choice([1, 3, 12, 100, 700]))
            ids.append(b.add_hull(rs.randn(n, 3) * 0.5, T[:3, :3], T[:3, 3], margin=m))
    host = b.build()
    pa = rs.randint(0, len(ids), 5000)
    pb = rs.randint(0, len(ids), 5000)
    _compare_gjk(host, pa, pb)


def test_intersection_kernel_matches_oracle(gpu):
    w = workloads.mixed_pairs(seed=6, n_pairs=20000)
    dev = w.scene.to_device()
    hit, st = kernels.gjk_intersection_batch(dev, w.pair_a, w.pair_b, want_status=True)
    ohit, _ = oracle.gjk_intersection(w.scene, w.pair_a, w.pair_b)
    assert np.array_equal(_np(hit).astype(np.int32), ohit)


def test_supports_bitwise(gpu):
    rs = np.random.RandomState(4)
    b = SceneBuilder()
    T = gb.random_pose(rs)
    b.add_sphere(T[:3, 3], 0.7)
    b.add_capsule(T, 0.3, 1.2)
    b.add_cylinder(T, 0.4, 0.9)
    b.add_ellipsoid(T, np.array([0.3, 0.6, 0.9]))
    b.add_box(T, np.array([0.3, 0.6, 0.9]))
    b.add_cone(T, 0.5, 1.1)
    b.add_disk(T[:3, 3], 0.6, T[:3, 2])
    b.add_ellipse(T[:3, 3], T[:3, :2].T, np.array([0.4, 0.8]))
    b.add_hull(rs.randn(10, 3), T[:3, :3], T[:3, 3])
    b.add_hull(rs.randn(3000, 3).astype(np.float32), T[:3, :3], T[:3, 3])
    b.add_hull(rs.randn(3000, 3))  # FP64 pool (values not FP32-exact)
    b.add_hull(rs.randn(500, 3), margin=0.2)
    host = b.build()
    n = host.n_shapes
    ids = np.repeat(np.arange(n), 200).astype(np.int32)
    dirs = rs.randn(len(ids), 3)
    dirs[::50] = 0.0  # zero-direction branches
    out = _np(kernels.support_batch(host.to_device(), ids, dirs))
    expect = oracle.support(host, ids, dirs)
    assert np.array_equal(out, expect)


def test_hull_argmax_ties_take_first_index(gpu):
    """Duplicate and coplanar vertices defeat the FP32 certificate; the FP64
    rescan must pick the first maximiser like np.argmax."""
    rs = np.random.RandomState(5)
    cube = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    v = np.vstack([cube] * 40)  # 320 vertices, each corner 40 times
    b = SceneBuilder()
    b.add_hull(v)
    host = b.build()
    dirs = np.vstack([np.eye(3), -np.eye(3), rs.randn(200, 3)])
    ids = np.zeros(len(dirs), dtype=np.int32)
    out = _np(kernels.support_batch(host.to_device(), ids, dirs))
    for d, p in zip(dirs, out):
        assert_array_almost_equal(p, v[np.argmax(v @ d)])
    assert np.array_equal(out, oracle.support(host, ids, dirs))


def test_tied_hull_vertices_every_lane_count(gpu):
    """Hulls whose support vertices tie exactly (duplicated corners, in one
    lane's stride or in several) drive the GJK sweeps' failed certificates
    through both rescans: the kept candidates of each lane when they cover the
    threshold, the full sweep when a lane holds three or more tied scores.
    Every lane count and layout must give the oracle's bits (np.argmax's first
    index on ties)."""
    from distance3d_amd.scene import DeviceScene
    rs = np.random.RandomState(11)
    cube = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    inner = rs.uniform(-0.5, 0.5, (100, 3))
    hulls = [np.vstack([cube] * 40),                     # every corner 40 times
             np.vstack([cube, inner, cube[::-1]]),        # every corner twice, far apart
             np.vstack([inner[:20], cube, cube]),          # twice, eight apart
             np.vstack([cube * 0.5, inner * 0.2, cube * 0.5 + 1e-9])]  # near ties
    b = SceneBuilder()
    hid = [b.add_hull(h) for h in hulls]
    partners = []
    for k in range(24):
        c = rs.uniform(-3.0, 3.0, 3)
        pose = np.eye(4)
        pose[:3, 3] = c
        partners.append(b.add_sphere(c, 0.3 + 0.5 * rs.rand()))
        partners.append(b.add_box(pose, rs.uniform(0.2, 1.5, 3)))
        partners.append(b.add_capsule(pose, 0.2, 1.0))
    host = b.build()
    pa = np.array([h for h in hid for _ in partners] + [hid[0], hid[1], hid[2]], dtype=np.int32)
    pb = np.array(partners * len(hid) + [hid[1], hid[2], hid[3]], dtype=np.int32)
    o = oracle.gjk_distance(host, pa, pb)
    for cmin in (0, 64):
        dev = DeviceScene(host, cluster_min=cmin)
        for g in (1, 2, 4, 8, 16, 32, 64):
            r = kernels.gjk_distance_batch(dev, pa, pb, lanes_per_pair=[1] + [g] * 15)
            assert np.array_equal(_np(r.status), o["status"]), (cmin, g)
            assert np.array_equal(_np(r.dist), o["dist"]), (cmin, g)
            assert np.array_equal(_np(r.cp_a), o["cp_a"], equal_nan=True), (cmin, g)
            assert np.array_equal(_np(r.cp_b), o["cp_b"], equal_nan=True), (cmin, g)


def test_aabbs_bitwise(gpu):
    w = workloads.mixed_pairs(seed=7, n_pairs=3000)
    ids = np.arange(w.scene.n_shapes, dtype=np.int32)
    out = _np(kernels.aabb_batch(w.scene.to_device(), ids))
    assert np.array_equal(out, oracle.aabbs(w.scene, ids))


def test_bad_shape_index_reports_status(gpu):
    w = workloads.hull_pairs(n_pairs=64)
    r = kernels.gjk_distance_batch(w.scene.to_device(), [0, 5, 10 ** 6], [1, -3, 2])
    st = _np(r.status)
    assert st[0] & 0xFF in (0, 1)
    assert st[1] == _native.BAD_INPUT and st[2] == _native.BAD_INPUT


def test_iteration_cap_status(gpu):
    w = workloads.hull_pairs(n_pairs=256)
    r = kernels.gjk_distance_batch(w.scene.to_device(), w.pair_a, w.pair_b, max_iterations=2)
    o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b, max_iterations=2)
    assert np.array_equal(_np(r.status), o["status"])
    assert np.any(o["status"] == _native.ITER_CAP)
    assert np.array_equal(_np(r.dist), o["dist"])


def test_simplex_output_matches_oracle(gpu):
    w = workloads.hull_pairs(n_pairs=512)
    r = kernels.gjk_distance_batch(w.scene.to_device(), w.pair_a, w.pair_b, want_simplex=True)
    o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b, want_simplex=True)
    assert np.array_equal(_np(r.n_points), o["n_points"])
    assert np.allclose(_np(r.simplex), o["simplex"], rtol=1e-9, atol=1e-12)


# ------------------------------------------ drop-in API, reference tests

def test_dropin_points_and_segments(gpu):  # test_gjk.py:7-53
    p1 = colliders.ConvexHullVertices(np.array([[0.0, 0.0, 0.0]]))
    p2 = colliders.ConvexHullVertices(np.array([[1.0, 0.0, 0.0]]))
    assert gjk.gjk_intersection(p1, p1)
    assert gjk.gjk(p1, p1)[0] == 0.0
    assert not gjk.gjk_intersection(p1, p2)
    assert gjk.gjk(p1, p2)[0] == 1.0
    s1 = colliders.ConvexHullVertices(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]))
    s2 = colliders.ConvexHullVertices(np.array([[0.0, 1.0, 0.0], [1.0, 1.0, 0.0]]))
    assert gjk.gjk(s1, s2)[0] == 1.0
    assert not gjk.gjk_intersection(s1, s2)
    s3 = colliders.ConvexHullVertices(np.array([[0.5, -1.0, 0.0], [0.5, 1.0, 0.0]]))
    assert gjk.gjk(s1, s3)[0] == 0.0
    assert gjk.gjk_intersection(s1, s3)


def test_dropin_boxes_known_distance(gpu):  # test_gjk.py:104-126
    T1 = np.array([[-0.29265666, -0.76990535, 0.56709596, 0.1867558],
                   [0.93923897, -0.12018753, 0.32153556, -0.09772779],
                   [-0.17939408, 0.62673815, 0.75829879, 0.09500884], [0., 0., 0., 1.]])
    T2 = T1.copy()
    T2[:3, 3] = [3.73511598, -1.95455576, 1.90017684]
    d = gjk.gjk(colliders.Box(T1, np.array([2.89098828, 1.15032456, 2.37517511])),
                colliders.Box(T2, np.array([0.96366276, 0.38344152, 0.79172504])))[0]
    assert d == approx(1.7900192730149391)


def test_dropin_capsules_closest_points(gpu):  # test_gjk.py:251-309
    c1 = colliders.Capsule(np.eye(4), 1, 1)
    T = np.eye(4)
    T[:3, 3] = [3, 0, 0]
    d, cp1, cp2, _ = gjk.gjk(c1, colliders.Capsule(T, 1, 1))
    assert d == approx(1)
    assert_array_almost_equal(cp1, [1, 0, -0.5])
    assert_array_almost_equal(cp2, [2, 0, -0.5])


def test_dropin_margin_and_too_far(gpu):  # test_gjk.py:476-482, test_gjk_jolt.py:88-95
    box = colliders.Box(np.eye(4), np.array([2.0, 2.0, 2.0]))
    sphere = colliders.Sphere(np.array([2.0, 0.0, 0.0]), 0.5)
    assert gjk.gjk(box, sphere)[0] == approx(gjk.gjk(colliders.Margin(box, 0.1), sphere)[0] + 0.1)
    far = gjk.gjk(colliders.ConvexHullVertices(np.zeros((1, 3))),
                  colliders.ConvexHullVertices(np.array([[1e8, 0.0, 0.0]])))
    assert far == (np.finfo(float).max, None, None, None)


def test_dropin_support_and_aabb(gpu):  # test_colliders.py:35-93
    convex = colliders.ConvexHullVertices(np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0]]))
    assert_array_almost_equal(convex.aabb(), [[0.0, 0.0], [0.0, 1.0], [0.0, 1.0]])
    cone = colliders.Cone(np.eye(4), 0.5, 1.0)
    assert_array_almost_equal(cone.aabb(), [[-0.5, 0.5], [-0.5, 0.5], [0.0, 1.0]])
    box = colliders.Margin(colliders.Box(np.eye(4), np.ones(3)), 0.1)
    assert_array_almost_equal(box.aabb(), [[-0.6, 0.6]] * 3)
    assert_array_almost_equal(colliders.Sphere(np.zeros(3), 1.0).support_function(
        np.array([0.0, 0.0, 0.0])), [0, 0, 1])


def test_cluster_layouts_and_group_sizes_agree_bitwise(gpu):
    """Pruned sweeps over the clustered layout give the oracle's supports for
    every lanes-per-pair value and clustering threshold (0 = plain layout)."""
    from distance3d_amd.scene import DeviceScene
    w = workloads.mixed_pairs(seed=12, n_pairs=12000)
    o = oracle.gjk_distance(w.scene, w.pair_a, w.pair_b)
    for cmin in (0, 16, 64):
        dev = DeviceScene(w.scene, cluster_min=cmin)
        # 8 lanes: slots whose cluster records leave no room for the survivor
        # list fall back to the per-lane scan
        for g in (1, 2, 4, 8, 16, 32, 64):
            r = kernels.gjk_distance_batch(dev, w.pair_a, w.pair_b,
                                           lanes_per_pair=[1] + [g] * 7)
            assert np.array_equal(_np(r.dist), o["dist"]), (cmin, g)
            assert np.array_equal(_np(r.status), o["status"]), (cmin, g)
            assert np.array_equal(_np(r.cp_a), o["cp_a"], equal_nan=True), (cmin, g)
