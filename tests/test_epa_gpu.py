"""GPU parity of the EPA kernel against the FP64 oracle's EPA.

Both sides are seeded with the same GJK simplices.  Tolerance: depth within
1e-6 relative, normals and contact points within 1e-6 (both evaluate the same
FP64 expressions; the polytope paths coincide), status identical.
"""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal
from pytest import approx

from distance3d_amd import colliders, epa, gjk, kernels, workloads
from distance3d_amd.scene import SceneBuilder
from oracle import oracle
from tests import geometry_brute as gb

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _np(t):
    return t.cpu().numpy()


def _epa_both(host, pa, pb):
    dev = host.to_device()
    g = kernels.gjk_distance_batch(dev, pa, pb, want_simplex=True)
    hit = _np(g.dist) == 0.0
    idx = np.nonzero(hit)[0]
    pa, pb = np.asarray(pa)[idx], np.asarray(pb)[idx]
    simplex, npts = _np(g.simplex)[idx], _np(g.n_points)[idx]
    r = kernels.epa_batch(dev, pa, pb, simplex, npts)
    o = oracle.epa(host, pa, pb, simplex, npts)
    return r, o, idx


def _check(r, o):
    st = _np(r.status)
    assert np.array_equal(st, o["status"])
    conv = st == 1
    d, od = _np(r.depth), o["depth"]
    assert np.all(np.abs(d - od) <= TOL * np.maximum(1.0, np.abs(od)))
    assert np.all(np.abs(_np(r.normal)[conv] - o["normal"][conv]) <= TOL)
    assert np.all(np.abs(_np(r.cp_a)[conv] - o["cp_a"][conv]) <= TOL * 10)
    assert np.all(np.abs(_np(r.cp_b)[conv] - o["cp_b"][conv]) <= TOL * 10)
    assert np.array_equal(_np(r.iterations), o["iterations"])
    return np.mean(d == od)


def test_epa_pairs_parity(gpu):
    w = workloads.epa_pairs(seed=9, n_pairs=4000)
    r, o, idx = _epa_both(w.scene, w.pair_a, w.pair_b)
    assert len(idx) > 3000
    same = _check(r, o)
    assert same > 0.99
    # converged depths never exceed the overlap the generator built in
    conv = _np(r.status) == 1
    assert np.all(_np(r.depth)[conv] <= w.meta["delta"][idx][conv] + 1e-9)
    # contact points reproduce the MTV
    assert_array_almost_equal((_np(r.cp_a) - _np(r.cp_b))[conv], _np(r.mtv)[conv], decimal=9)


def test_epa_mixed_primitives_parity(gpu):
    rs = np.random.RandomState(31)
    b = SceneBuilder()
    for k in range(300):
        T = gb.random_pose(rs, 0.5)
        kind = k % 6
        if kind == 0:
            b.add_sphere(T[:3, 3], rs.uniform(0.2, 1))
        elif kind == 1:
            b.add_capsule(T, rs.uniform(0.1, 0.5), rs.uniform(0.2, 2))
        elif kind == 2:
            b.add_cylinder(T, rs.uniform(0.1, 0.5), rs.uniform(0.2, 2))
        elif kind == 3:
            b.add_box(T, rs.uniform(0.2, 1.5, 3))
        elif kind == 4:
            b.add_ellipsoid(T, rs.uniform(0.2, 1, 3))
        else:
            b.add_hull(rs.randn(int(rs.choice([5, 40, 300])), 3) * 0.6, T[:3, :3], T[:3, 3])
    host = b.build()
    pa = rs.randint(0, 300, 3000)
    pb = rs.randint(0, 300, 3000)
    r, o, idx = _epa_both(host, pa, pb)
    assert len(idx) > 500
    _check(r, o)


def test_epa_tied_hull_vertices_bitwise(gpu):
    """Hulls whose support vertices tie exactly (duplicated corners) make EPA's
    cached sweeps fail their certificate along many face normals; both
    rescans (the lanes' kept candidates, or the full sweep) must reproduce
    the oracle's first-index choices bit for bit."""
    rs = np.random.RandomState(13)
    cube = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    inner = rs.uniform(-0.5, 0.5, (100, 3))
    b = SceneBuilder()
    hid = [b.add_hull(np.vstack([cube] * 40)),
           b.add_hull(np.vstack([cube, inner, cube[::-1]])),
           b.add_hull(np.vstack([inner[:20], cube, cube]) * 0.8)]
    partners = []
    for k in range(40):
        T = gb.random_pose(rs, 0.8)
        partners.append(b.add_box(T, rs.uniform(0.5, 1.5, 3)))
        partners.append(b.add_sphere(T[:3, 3], rs.uniform(0.3, 0.9)))
    host = b.build()
    pa = np.array([h for h in hid for _ in partners] + [hid[0], hid[0], hid[1]], dtype=np.int32)
    pb = np.array(partners * len(hid) + [hid[1], hid[2], hid[2]], dtype=np.int32)
    r, o, idx = _epa_both(host, pa, pb)
    assert len(idx) > 100
    assert np.array_equal(_np(r.status), o["status"])
    assert np.array_equal(_np(r.depth), o["depth"])
    assert np.array_equal(_np(r.normal), o["normal"])
    assert np.array_equal(_np(r.cp_a), o["cp_a"])
    assert np.array_equal(_np(r.iterations), o["iterations"])


def test_epa_small_hulls_vs_minkowski_brute_force(gpu):
    rs = np.random.RandomState(26)
    b = SceneBuilder()
    verts = []
    for _ in range(80):
        v1 = rs.randn(rs.randint(5, 14), 3)
        v2 = rs.randn(rs.randint(5, 14), 3) * 0.8 + rs.randn(3) * 0.5
        b.add_hull(v1)
        b.add_hull(v2)
        verts.append((v1, v2))
    host = b.build()
    pa, pb = np.arange(0, 160, 2), np.arange(1, 160, 2)
    r, o, idx = _epa_both(host, pa, pb)
    done = 0
    for j, k in enumerate(idx):
        if _np(r.status)[j] != 1:
            continue
        depth, _ = gb.penetration_depth(*verts[k])
        assert _np(r.depth)[j] == approx(depth, rel=1e-7, abs=1e-9)
        done += 1
    assert done > 40


def test_dropin_epa_known_mtv(gpu):  # test_epa.py:7-34
    v1 = np.array([[1.76405235, 0.40015721, 0.97873798], [2.2408932, 1.86755799, -0.97727788],
                   [0.4105985, 0.14404357, 1.45427351], [0.33367433, 1.49407907, -0.20515826],
                   [0.3130677, -0.85409574, -2.55298982], [2.26975462, -1.45436567, 0.04575852],
                   [-0.18718385, 1.53277921, 1.46935877]])
    v2 = np.array([[-2.32605299, -0.31242692, 0.07599278], [0.88503416, 1.23786508, -0.467683],
                   [-0.64755927, -1.01306774, -1.50037412], [-2.05152671, 1.98626062, -0.59000837],
                   [-0.78333082, -1.21731013, 0.69713417], [-1.95915437, -0.17725505, -0.97582275],
                   [0.04164598, -0.47531991, -1.26098837], [-0.37343875, 0.4638171, -0.01383896],
                   [-0.04278462, -0.59883687, -0.44309735]])
    c1, c2 = colliders.ConvexHullVertices(v1), colliders.ConvexHullVertices(v2)
    dist, p1, p2, simplex = gjk.gjk(c1, c2)
    assert_array_almost_equal(p1, p2)
    mtv, _, success = epa.epa(simplex, c1, c2)
    assert success
    assert_array_almost_equal(mtv, np.array([-0.387287, 0.179576, -0.176204]))
    # a plain (4, 3) array works too
    mtv2, _, success2 = epa.epa(np.asarray(simplex).copy(), c1, c2)
    assert success2
    assert_array_almost_equal(mtv2, mtv)


def test_dropin_epa_separation(gpu):  # test_epa.py:37-60
    rs = np.random.RandomState(25)
    done = 0
    for _ in range(50):
        v1 = rs.rand(6, 3) * np.array([[2, 2, 2]])
        v2 = rs.rand(6, 3) * np.array([[1, 1, 1]])
        c1, c2 = colliders.ConvexHullVertices(v1), colliders.ConvexHullVertices(v2)
        dist, _, _, simplex = gjk.gjk(c1, c2)
        if dist > 0.0:
            continue
        mtv, _, success = epa.epa(simplex, c1, c2)
        assert success
        depth = np.linalg.norm(mtv)
        pdir = mtv / depth
        moved = colliders.ConvexHullVertices(v2 + (1.0 + depth) * pdir[np.newaxis])
        assert gjk.gjk(c1, moved)[0] == approx(1.0)
        done += 1
    assert done > 10


def test_epa_contact_points_on_both_hulls(gpu):
    """GPU contact points lie on the boundary of their own hull (largest
    Qhull facet offset zero) and reproduce the MTV, on random hull pairs."""
    from scipy.spatial import ConvexHull
    rs = np.random.RandomState(26)
    b = SceneBuilder()
    verts = []
    for _ in range(120):
        v1 = rs.randn(rs.randint(5, 14), 3)
        v2 = rs.randn(rs.randint(5, 14), 3) * 0.8 + rs.randn(3) * 0.5
        b.add_hull(v1)
        b.add_hull(v2)
        verts += [v1, v2]
    host = b.build()
    pa = np.arange(0, 240, 2, dtype=np.int32)
    pb = pa + 1
    r, o, idx = _epa_both(host, pa, pb)
    st = _np(r.status)
    done = 0
    for k, i in enumerate(idx):
        if st[k] != 1:
            continue
        for v, cp in ((verts[2 * i], _np(r.cp_a)[k]), (verts[2 * i + 1], _np(r.cp_b)[k])):
            eq = ConvexHull(v).equations
            assert abs((eq[:, :3] @ cp + eq[:, 3]).max()) < 1e-9
        assert_array_almost_equal(_np(r.cp_a)[k] - _np(r.cp_b)[k], _np(r.mtv)[k], decimal=12)
        done += 1
    assert done > 40


def test_epa_faces_match_oracle(gpu):
    """The final polytope (faces output) equals the oracle's, face for face
    and bit for bit, in the same order."""
    w = workloads.epa_pairs(seed=12, n_pairs=600)
    dev = w.scene.to_device()
    g = kernels.gjk_distance_batch(dev, w.pair_a, w.pair_b, want_simplex=True)
    idx = np.nonzero(_np(g.dist) == 0.0)[0]
    pa, pb = np.asarray(w.pair_a)[idx], np.asarray(w.pair_b)[idx]
    simplex, npts = _np(g.simplex)[idx], _np(g.n_points)[idx]
    r = kernels.epa_batch(dev, pa, pb, simplex, npts, want_faces=True)
    o = oracle.epa(w.scene, pa, pb, simplex, npts, want_faces=True)
    assert np.array_equal(_np(r.status), o["status"])
    nf = _np(r.n_faces)
    assert np.array_equal(nf, o["n_faces"])
    F = _np(r.faces)
    for k in range(len(idx)):
        assert np.array_equal(F[k, :nf[k]], o["faces"][k, :nf[k]])


def test_epa_dropin_faces_and_caller_simplex(gpu):  # test_epa.py:7-34
    """epa.epa returns the polytope (n_faces, 4, 3) with unit normals whose
    closest face lies within epsilon of the depth; a plain (4, 3) simplex
    whose rows span a tetrahedron seeds the polytope as given, with the
    oracle's result from the same rows."""
    from tests.test_oracle_pins import _closed_convex_polytope
    v1 = np.array([[1.76405235, 0.40015721, 0.97873798], [2.2408932, 1.86755799, -0.97727788],
                   [0.4105985, 0.14404357, 1.45427351], [0.33367433, 1.49407907, -0.20515826],
                   [0.3130677, -0.85409574, -2.55298982], [2.26975462, -1.45436567, 0.04575852],
                   [-0.18718385, 1.53277921, 1.46935877]])
    v2 = np.array([[-2.32605299, -0.31242692, 0.07599278], [0.88503416, 1.23786508, -0.467683],
                   [-0.64755927, -1.01306774, -1.50037412], [-2.05152671, 1.98626062, -0.59000837],
                   [-0.78333082, -1.21731013, 0.69713417], [-1.95915437, -0.17725505, -0.97582275],
                   [0.04164598, -0.47531991, -1.26098837], [-0.37343875, 0.4638171, -0.01383896],
                   [-0.04278462, -0.59883687, -0.44309735]])
    c1, c2 = colliders.ConvexHullVertices(v1), colliders.ConvexHullVertices(v2)
    dist, p1, p2, simplex = gjk.gjk(c1, c2)
    assert dist == 0.0
    mtv, faces, success = epa.epa(simplex, c1, c2)
    assert success
    assert_array_almost_equal(mtv, [-0.387287, 0.179576, -0.176204])
    assert faces.ndim == 3 and faces.shape[1:] == (4, 3) and len(faces) >= 4
    md = (v1[:, None, :] - v2[None, :, :]).reshape(-1, 3)
    _closed_convex_polytope(faces, np.linalg.norm(mtv), md)
    # a caller-built simplex: four Minkowski-difference points spanning a
    # tetrahedron around the origin, passed as a plain array
    b = SceneBuilder()
    b.add_hull(v1)
    b.add_hull(v2)
    host = b.build()
    Y = np.asarray(simplex, dtype=np.float64)[:4] * 0.5
    if simplex.n_points < 4:
        Y = 1e-3 * np.array([[1.0, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])
    mtv2, faces2, ok2 = epa.epa(np.array(Y), c1, c2)
    seed = np.stack([Y, Y, np.zeros((4, 3))])[None]
    o = oracle.epa(host, [0], [1], seed, [4], max_faces=64, want_faces=True)
    assert ok2 == (o["status"][0] == 1)
    assert np.array_equal(mtv2, o["mtv"][0])
    assert np.array_equal(faces2, o["faces"][0, :o["n_faces"][0]])
    assert_array_almost_equal(mtv2, [-0.387287, 0.179576, -0.176204])
    # max_iter past 64 is accepted and converges to the same answer
    mtv3, _, ok3 = epa.epa(simplex, c1, c2, max_iter=500)
    assert ok3 and np.array_equal(mtv3, mtv)


def test_epa_long_stalled_run_matches_oracle(gpu):
    """epsilon = 0 and 400 iterations from a tiny caller-built seed: the
    kernel recycles its 72 vertex rows many times (the oracle at 260), with
    the oracle's status, depth, normal, contact points and polytope."""
    from tests.test_oracle_pins import stalled_epa_case
    host, pa, pb, simplex, npts = stalled_epa_case()
    dev = host.to_device()
    r = kernels.epa_batch(dev, pa, pb, simplex, npts, max_iterations=400, epsilon=0.0,
                          want_faces=True)
    o = oracle.epa(host, pa, pb, simplex, npts, max_iterations=400, epsilon=0.0,
                   want_faces=True)
    assert np.array_equal(_np(r.status), o["status"])
    assert np.array_equal(_np(r.iterations), o["iterations"])
    assert np.sum(o["iterations"] == 400) >= 5
    assert np.array_equal(_np(r.raw), np.concatenate(
        [o["depth"][:, None], o["normal"], o["mtv"], o["cp_a"], o["cp_b"]], axis=1))
    nf = _np(r.n_faces)
    assert np.array_equal(nf, o["n_faces"])
    F = _np(r.faces)
    for k in range(len(pa)):
        assert np.array_equal(F[k, :nf[k]], o["faces"][k, :nf[k]])
