"""GPU parity on degenerate geometry: flat, collinear and single-point hulls,
hulls of a few repeated vertices, zero-radius spheres and zero-height
capsules, boxes in exact face / edge / corner contact and one 2^-40 gap
apart, a collider against itself, and tiny shapes far from the origin.
These drive the simplex solver's degenerate branches (zero-area triangles,
zero-volume tetrahedra, coincident support points) and the hull sweeps' ties,
where any difference in operation order would show.

Bar: the GJK outputs (distance, closest points, status, iteration count)
and the intersection flags equal the FP64 oracle's bit for bit for every
lanes-per-pair value; EPA on the intersecting pairs gives the oracle's
status, iteration count, depth and normal bit for bit; MPR its status and
penetration outputs bit for bit.
"""
import numpy as np
import pytest

from distance3d_amd import kernels
from distance3d_amd.scene import SceneBuilder
from oracle import oracle
from tests import geometry_brute as gb


def _np(t):
    return t.cpu().numpy()


def degenerate_scene(seed=0):
    rs = np.random.RandomState(seed)
    b = SceneBuilder()
    ids = []

    def pose(scale=0.8):
        return gb.random_pose(rs, scale)

    def hull(v):
        T = pose()
        ids.append(b.add_hull(np.asarray(v, dtype=np.float32), T[:3, :3], T[:3, 3]))

    for n in (3, 4, 20, 200, 1500):  # flat: every vertex in one plane
        for _ in range(6):
            v = np.zeros((n, 3))
            v[:, :2] = rs.randn(n, 2) * 0.6
            hull(v)
    for n in (2, 5, 50):  # collinear
        for _ in range(6):
            hull(np.outer(rs.randn(n), rs.randn(3)) * 0.5)
    for n in (1, 8, 100):  # one point, repeated
        for _ in range(4):
            hull(np.repeat(rs.randn(1, 3) * 0.3, n, axis=0))
    for _ in range(6):  # 300 vertices drawn from 6 distinct points
        hull((rs.randn(6, 3) * 0.5)[rs.randint(0, 6, 300)])
    for _ in range(6):  # point-like primitives
        T = pose()
        ids.append(b.add_sphere(T[:3, 3], 0.0))
        ids.append(b.add_capsule(T, rs.uniform(0.1, 0.4), 0.0))
    # unit boxes in face, edge and corner contact, a 2^-40 gap apart, and coincident
    boxes = []
    for off in ([1.0, 0, 0], [1.0, 1.0, 0], [1.0, 1.0, 1.0], [1.0 + 2.0 ** -40, 0, 0],
                [0.0, 0, 0]):
        T1 = np.eye(4)
        T1[:3, 3] = off
        boxes.append((b.add_box(np.eye(4), [1.0, 1.0, 1.0]), b.add_box(T1, [1.0, 1.0, 1.0])))
    far = []  # sub-millimetre shapes about 2,300 units from the origin
    for _ in range(12):
        T = pose(1e-4)
        T[:3, 3] += [1e3, -2e3, 5e2]
        far.append(b.add_box(T, rs.uniform(1e-4, 3e-4, 3)))
        T = pose(1e-4)
        T[:3, 3] += [1e3, -2e3, 5e2]
        far.append(b.add_hull((rs.randn(30, 3) * 1e-4).astype(np.float32), T[:3, :3], T[:3, 3]))
    host = b.build()
    pa = list(rs.randint(0, len(ids), 3000))
    pb = list(rs.randint(0, len(ids), 3000))
    pa += ids[:40]  # a collider against itself
    pb += ids[:40]
    pa += [i for i, _ in boxes]
    pb += [j for _, j in boxes]
    pa += [far[k] for k in rs.randint(0, len(far), 300)]
    pb += [far[k] for k in rs.randint(0, len(far), 300)]
    return host, np.array(pa, np.int32), np.array(pb, np.int32)


def test_degenerate_scene_oracle_is_finite():
    """CPU: the oracle's outputs on the scene are finite (so the GPU
    comparison below compares numbers), and the scene reaches the cases it is
    meant to: touching boxes intersect, EPA sees every status but the
    iteration cap."""
    host, pa, pb = degenerate_scene()
    g = oracle.gjk_distance(host, pa, pb, want_simplex=True)
    assert np.isfinite(g["dist"]).all() and np.isfinite(g["cp_a"]).all()
    assert np.isfinite(g["cp_b"]).all()
    assert np.all(g["dist"][3040:3045] == 0.0)
    hit = np.nonzero(g["dist"] == 0.0)[0]
    assert len(hit) > 400
    e = oracle.epa(host, pa[hit], pb[hit], g["simplex"][hit], g["n_points"][hit])
    assert np.isfinite(e["depth"]).all()
    assert set(np.unique(e["status"])) == {0, 1, 2, 3}


@pytest.mark.gpu
def test_degenerate_geometry_gjk_bitwise(gpu):
    host, pa, pb = degenerate_scene()
    dev = host.to_device()
    o = oracle.gjk_distance(host, pa, pb)
    for lanes in (None, [1] + [4] * 7, [1] + [16] * 7, [1] + [64] * 7):
        kw = {} if lanes is None else {"lanes_per_pair": lanes}
        r = kernels.gjk_distance_batch(dev, pa, pb, **kw)
        assert np.array_equal(_np(r.status), o["status"]), lanes
        assert np.array_equal(_np(r.iterations), o["iterations"]), lanes
        assert np.array_equal(_np(r.dist), o["dist"]), lanes
        assert np.array_equal(_np(r.cp_a), o["cp_a"]), lanes
        assert np.array_equal(_np(r.cp_b), o["cp_b"]), lanes
    hit, _ = kernels.gjk_intersection_batch(dev, pa, pb, want_status=True)
    ohit, _ = oracle.gjk_intersection(host, pa, pb)
    assert np.array_equal(_np(hit).astype(np.int32), ohit)


@pytest.mark.gpu
def test_degenerate_geometry_epa_bitwise(gpu):
    host, pa, pb = degenerate_scene()
    dev = host.to_device()
    g = kernels.gjk_distance_batch(dev, pa, pb, want_simplex=True)
    idx = np.nonzero(_np(g.dist) == 0.0)[0]
    simplex, npts = _np(g.simplex)[idx], _np(g.n_points)[idx]
    r = kernels.epa_batch(dev, pa[idx], pb[idx], simplex, npts)
    o = oracle.epa(host, pa[idx], pb[idx], simplex, npts)
    assert np.array_equal(_np(r.status), o["status"])
    assert np.array_equal(_np(r.iterations), o["iterations"])
    assert np.array_equal(_np(r.depth), o["depth"])
    conv = o["status"] == 1
    assert np.array_equal(_np(r.normal)[conv], o["normal"][conv])


@pytest.mark.gpu
def test_degenerate_geometry_mpr_bitwise(gpu):
    """MPR from the colliders' centres (coincident for a collider against
    itself): status and, with penetration info, depth, direction and contact
    bit for bit."""
    from distance3d_amd.scene import collider_centers
    host, pa, pb = degenerate_scene()
    centers = collider_centers(host)
    dev = host.to_device()
    for penetration in (True, False):
        st, out = kernels.mpr_batch(dev, pa, pb, centers, penetration=penetration)
        est, eout = oracle.mpr(host, pa, pb, centers, penetration=penetration)
        assert np.array_equal(st.cpu().numpy(), est)
        if penetration:
            assert np.array_equal(out.cpu().numpy(), eout, equal_nan=True)
            assert np.isfinite(eout[est == 1]).all() and np.sum(est == 1) > 400


@pytest.mark.gpu
def test_degenerate_geometry_aabbs_and_overlaps_exact(gpu):
    """AABBs of the degenerate colliders (zero extent along some axes) bit for
    bit, and the overlap pairs among them, touching boxes included, equal to
    the brute-force closed-interval set."""
    host, _, _ = degenerate_scene()
    ids = np.arange(host.n_shapes, dtype=np.int32)
    aabbs = kernels.aabb_batch(host.to_device(), ids)
    host_aabbs = oracle.aabbs(host, ids)
    assert np.array_equal(aabbs.cpu().numpy(), host_aabbs)
    bvh = kernels.bvh_build(aabbs)
    g = kernels.bvh_query(bvh, aabbs).cpu().numpy().astype(np.int64)
    o, cnt = oracle.all_aabbs_overlap(host_aabbs, host_aabbs)
    assert len(o) == cnt
    o = o[np.lexsort((o[:, 1], o[:, 0]))]
    assert np.array_equal(g, o)
