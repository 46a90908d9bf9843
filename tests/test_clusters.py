"""Clustered device layout of hull vertex sets (d3d_cluster_scene, host code):
the layout is a per-hull permutation that keeps records sharing a range
together, clustered hulls are aligned, every cluster's sphere contains its
vertices, and the FP32 bounds the device evaluates (c.d +- r|d| as one FMA
chain, |d| rounded up) bracket the exact score of every cluster vertex."""
import numpy as np
import pytest

from distance3d_amd import _native
from distance3d_amd.scene import SceneBuilder, cluster_layout, HostScene

K = _native.CLUSTER_SIZE


def _scene(rng):
    b = SceneBuilder()
    for n in (5, 16, 17, 31, 32, 33, 64, 100, 257, 1000):
        v = rng.normal(size=(n, 3)) * rng.uniform(0.1, 3.0, 3)
        if n % 2:
            v = v.astype(np.float32).astype(np.float64)  # FP32-exact: no FP64 copy
        b.add_hull(v)
        b.add_sphere(rng.normal(size=3), 0.5)
    # a hull with duplicated vertices (exact ties)
    v = rng.normal(size=(40, 3)).astype(np.float32).astype(np.float64)
    b.add_hull(np.vstack([v, v[::-1]]))
    host = b.build()
    # a second record sharing the vertex range of record 0
    f = np.vstack([host.shape_f64, host.shape_f64[:1]])
    i = np.vstack([host.shape_i32, host.shape_i32[:1]])
    return HostScene(f, i, host.vx, host.vy, host.vz, host.v64)


def _fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def test_layout_is_a_per_hull_permutation():
    rng = np.random.default_rng(0)
    host = _scene(rng)
    lay = cluster_layout(host, min_vertices=32)
    hi, li = host.shape_i32, lay["shape_i32"]
    assert np.array_equal(hi[:, :2], li[:, :2])
    assert np.array_equal(np.delete(host.shape_f64, [13, 14], axis=1),
                          np.delete(lay["shape_f64"], [13, 14], axis=1))
    for k in np.nonzero(hi[:, 0] == _native.HULL)[0]:
        n, o, o64 = hi[k, 1], hi[k, 2], hi[k, 3]
        no, no64 = li[k, 2], li[k, 3]
        perm = lay["vorig"][no:no + n]
        assert np.array_equal(np.sort(perm), np.arange(n))
        for src, dst in (("vx", "vx"), ("vy", "vy"), ("vz", "vz")):
            assert np.array_equal(lay[dst][no:no + n], getattr(host, src)[o:o + n][perm])
        assert (o64 < 0) == (no64 < 0)
        if o64 >= 0:
            assert np.array_equal(lay["v64"][no64:no64 + n], host.v64[o64:o64 + n][perm])
        nc = lay["shape_f64"][k, 14]
        if n >= 32:
            assert no % K == 0 and nc == -(-n // K)
        else:
            assert nc == 0 and np.array_equal(perm, np.arange(n))
    # records sharing a range keep sharing it
    assert np.array_equal(li[0], li[-1]) and np.array_equal(lay["shape_f64"][0], lay["shape_f64"][-1])


def test_cluster_bound_never_below_exact_scores():
    rng = np.random.default_rng(1)
    host = _scene(rng)
    lay = cluster_layout(host, min_vertices=32)
    li, f = lay["shape_i32"], lay["shape_f64"]
    dirs = rng.normal(size=(200, 3)) * rng.uniform(1e-3, 1e3, (200, 1))
    dirs = np.vstack([dirs, np.eye(3), -np.eye(3)])
    checked = 0
    for k in np.nonzero((li[:, 0] == _native.HULL) & (f[:, 14] > 0))[0]:
        n, no, no64 = li[k, 1], li[k, 2], li[k, 3]
        c0, nc = int(f[k, 13]), int(f[k, 14])
        if no64 >= 0:
            V = lay["v64"][no64:no64 + n].astype(np.longdouble)
        else:
            V = np.stack([lay["vx"][no:no + n], lay["vy"][no:no + n],
                          lay["vz"][no:no + n]], 1).astype(np.longdouble)
        C = lay["clusters"][c0:c0 + nc]
        for l in dirs:
            d = l.astype(np.float32)
            nl = np.float32(np.sqrt(l @ l) * (1.0 + 2.0 ** -19))
            dot = _fma32(C[:, 2], d[2], _fma32(C[:, 1], d[1], (C[:, 0] * d[0]).astype(np.float32)))
            ub = _fma32(C[:, 3], np.float32(nl), dot)
            lb = _fma32(-C[:, 3], np.float32(nl), dot)
            exact = V @ l.astype(np.longdouble)
            per_cluster = np.array([exact[q * K:(q + 1) * K].max() for q in range(nc)])
            per_cluster_min = np.array([exact[q * K:(q + 1) * K].min() for q in range(nc)])
            assert np.all(ub.astype(np.longdouble) >= per_cluster)
            # c.d - r|d| with the same inflation is never above a vertex score
            assert np.all(lb.astype(np.longdouble) <= per_cluster_min)
            checked += 1
        # containment with room to spare for the inflation
        for q in range(nc):
            P = V[q * K:(q + 1) * K]
            r = np.sqrt(((P - C[q, :3].astype(np.longdouble)) ** 2).sum(1)).max()
            assert r <= C[q, 3]
    assert checked > 0


def test_non_finite_hull_keeps_order_and_unbounded_clusters():
    v = np.random.default_rng(2).normal(size=(40, 3))
    v[7, 1] = np.inf
    b = SceneBuilder()
    b.add_hull(v)
    lay = cluster_layout(b.build(), min_vertices=16)
    assert np.array_equal(lay["vorig"][:40], np.arange(40))
    assert np.all(np.isinf(lay["clusters"][:, 3]))


def test_bad_record_is_rejected():
    b = SceneBuilder()
    b.add_hull(np.eye(3))
    host = b.build()
    host.shape_i32[0, 0] = 42
    with pytest.raises(_native.D3DError):
        cluster_layout(host)
