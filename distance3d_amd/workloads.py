"""Synthetic collision workloads, packed straight into scene arrays.

Every generator uses ``numpy.random.default_rng(seed)``; rotations are
uniform quaternions (normalised N(0, I4)).  Hull vertices are drawn in FP64,
rounded once to FP32 in the collider's local frame, and that FP32 set with an
FP64 pose is the geometry both the GPU and the FP64 checker see.

* ``hull_pairs``   -- 4096 pairs of 32-vertex hulls (unit-sphere points scaled
  by U(0.2, 1), centres U[-2, 2]^3), seed 0.
* ``mixed_pairs``  -- 1,048,576 pairs, each side uniform over {sphere, capsule,
  cylinder, box, ellipsoid, hull with log-uniform 8..4096 vertices on a
  perturbed unit sphere}, centres U[-3, 3]^3, seed 1.
* ``epa_pairs``    -- 262,144 overlapping hull pairs, 64..256 vertices on
  ellipsoid surfaces (radii U(0.3, 1)), projections along a random axis
  overlapping by U(0.01, 0.5), seed 2.
* ``broad_phase_scene`` -- 1,048,576 oriented boxes (sizes log-uniform
  0.01..0.5 per axis) and 65,536 capsules (r U(0.05, 0.5), h U(0.1, 2)) with
  centres uniform in [0, 100]^3, seed 3; boxes are shapes 0..n-1, capsules
  follow.
* ``serving_workload`` -- a resident scene of 512 colliders (1/3 capsules, 1/3
  boxes, 1/3 hulls with 64..1024 vertices), a table of random poses in a
  [0, 4]^3 workspace, and a Poisson stream of requests (mean rate 20,000/s)
  of log-uniform 1..4096 pairs, each pair with its own pose rows, seed 4.
"""
import numpy as np

from . import _native
from .scene import F64_COLS, I32_COLS, HostScene

_ULP_MARGIN = 1.0 + 2.0 ** -20


def uniform_rotations(rng, n):
    """(n, 3, 3) rotation matrices from uniform unit quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _bounds_per_hull(v32, counts):
    """max(|x|+|y|+|z|) per hull over a concatenated (N,3) FP32 array."""
    s = np.abs(v32.astype(np.float64)).sum(axis=1)
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    return np.maximum.reduceat(s, starts) * _ULP_MARGIN


class Workload:
    """A packed scene plus the pair list to evaluate on it."""

    def __init__(self, name, scene, pair_a, pair_b, meta=None):
        self.name = name
        self.scene = scene
        self.pair_a = np.ascontiguousarray(pair_a, dtype=np.int32)
        self.pair_b = np.ascontiguousarray(pair_b, dtype=np.int32)
        self.meta = meta or {}

    @property
    def n_pairs(self):
        return len(self.pair_a)

    def hull_bytes_per_pair(self):
        """12 * (n_A + n_B): FP32 hull bytes a pair reads once."""
        i = self.scene.shape_i32
        nv = np.where(i[:, 0] == _native.HULL, i[:, 1], 0).astype(np.int64)
        return 12 * (nv[self.pair_a] + nv[self.pair_b])

    def subset(self, idx):
        idx = np.asarray(idx)
        return Workload(self.name + "[subset]", self.scene, self.pair_a[idx], self.pair_b[idx],
                        dict(self.meta))


def _scene_from(types, R, t, params, hull_counts, v32):
    n = len(types)
    f = np.zeros((n, F64_COLS))
    f[:, 0:9] = R.reshape(n, 9)
    f[:, 9:12] = t
    f[:, 12:15] = params
    i = np.zeros((n, I32_COLS), dtype=np.int32)
    i[:, 0] = types
    i[:, 3] = -1
    hull = types == _native.HULL
    counts = hull_counts[hull].astype(np.int64)
    offs = np.concatenate(([0], np.cumsum(counts)[:-1]))
    i[hull, 1] = counts
    i[hull, 2] = offs
    if len(v32):
        f[hull, 12] = _bounds_per_hull(v32, counts)
        vx, vy, vz = (np.ascontiguousarray(v32[:, k]) for k in range(3))
    else:
        vx = vy = vz = np.zeros(0, dtype=np.float32)
    return HostScene(f, i, vx, vy, vz)


def _sphere_points(rng, total, chunk=1 << 24):
    """(total, 3) FP64 unit vectors, generated in chunks."""
    out = np.empty((total, 3))
    for s in range(0, total, chunk):
        e = min(total, s + chunk)
        g = rng.standard_normal((e - s, 3))
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        out[s:e] = g
    return out


def hull_pairs(seed=0, n_pairs=4096, n_vertices=32):
    """4096 pairs of 32-vertex hulls (the CPU-runnable configuration)."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    v = _sphere_points(rng, n * n_vertices).reshape(n, n_vertices, 3)
    s = rng.uniform(0.2, 1.0, n)
    R = uniform_rotations(rng, n)
    c = rng.uniform(-2.0, 2.0, (n, 3))
    v32 = (v * s[:, None, None]).astype(np.float32).reshape(-1, 3)
    types = np.full(n, _native.HULL, dtype=np.int32)
    counts = np.full(n, n_vertices, dtype=np.int64)
    scene = _scene_from(types, R, c, np.zeros((n, 3)), counts, v32)
    return Workload("hull_pairs", scene, np.arange(0, n, 2), np.arange(1, n, 2),
                    {"seed": seed, "n_pairs": n_pairs, "n_vertices": n_vertices})


MIXED_TYPES = np.array([_native.SPHERE, _native.CAPSULE, _native.CYLINDER, _native.BOX,
                        _native.ELLIPSOID, _native.HULL], dtype=np.int32)


def mixed_pairs(seed=1, n_pairs=1 << 20, min_vertices=8, max_vertices=4096):
    """Mixed-primitive pairs; every pair has its own two colliders."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    types = MIXED_TYPES[rng.integers(0, len(MIXED_TYPES), n)]
    R = uniform_rotations(rng, n)
    t = rng.uniform(-3.0, 3.0, (n, 3))
    params = np.zeros((n, 3))
    u = rng.uniform(size=(n, 3))
    sph = types == _native.SPHERE
    params[sph, 0] = 0.1 + 0.9 * u[sph, 0]
    R[sph] = np.eye(3)
    for ty in (_native.CAPSULE, _native.CYLINDER):
        m = types == ty
        params[m, 0] = 0.05 + 0.45 * u[m, 0]
        params[m, 1] = 0.2 + 1.8 * u[m, 1]
    box = types == _native.BOX
    params[box] = 0.1 + 1.9 * u[box]
    ell = types == _native.ELLIPSOID
    params[ell] = 0.1 + 0.9 * u[ell]
    hull = types == _native.HULL
    counts = np.zeros(n, dtype=np.int64)
    nh = int(hull.sum())
    counts[hull] = np.rint(np.exp(rng.uniform(np.log(min_vertices), np.log(max_vertices), nh)))
    total = int(counts.sum())
    v = _sphere_points(rng, total)
    radius = 1.0 + rng.uniform(-0.1, 0.1, total)
    v32 = (v * radius[:, None]).astype(np.float32)
    del v, radius
    scene = _scene_from(types, R, t, params, counts, v32)
    return Workload("mixed_pairs", scene, np.arange(0, n, 2), np.arange(1, n, 2),
                    {"seed": seed, "n_pairs": n_pairs})


def epa_pairs(seed=2, n_pairs=262144, min_vertices=64, max_vertices=256):
    """Overlapping hull pairs for penetration depth."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    counts = rng.integers(min_vertices, max_vertices + 1, n).astype(np.int64)
    radii = rng.uniform(0.3, 1.0, (n, 3))
    total = int(counts.sum())
    owner = np.repeat(np.arange(n), counts)
    v = _sphere_points(rng, total) * radii[owner]
    v32 = v.astype(np.float32)
    del v
    R = uniform_rotations(rng, n)
    axis = _sphere_points(rng, n_pairs)
    delta = rng.uniform(0.01, 0.5, n_pairs)
    cA = rng.uniform(-3.0, 3.0, (n_pairs, 3))
    # support values h_A(u) and h_B(-u) of the posed FP32 hulls (centred at 0)
    local_dir = np.empty((n, 3))
    local_dir[0::2] = np.einsum("kji,kj->ki", R[0::2], axis)
    local_dir[1::2] = np.einsum("kji,kj->ki", R[1::2], -axis)
    proj = np.einsum("ij,ij->i", v32.astype(np.float64), local_dir[owner])
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    h = np.maximum.reduceat(proj, starts)
    gap = h[0::2] + h[1::2] - delta
    cB = cA + axis * gap[:, None]
    t = np.empty((n, 3))
    t[0::2] = cA
    t[1::2] = cB
    types = np.full(n, _native.HULL, dtype=np.int32)
    scene = _scene_from(types, R, t, np.zeros((n, 3)), counts, v32)
    return Workload("epa_pairs", scene, np.arange(0, n, 2), np.arange(1, n, 2),
                    {"seed": seed, "n_pairs": n_pairs, "delta": delta, "axis": axis})


def broad_phase_scene(seed=3, n_boxes=1 << 20, n_capsules=65536, extent=100.0):
    """Boxes and query capsules for the AABB broad phase."""
    rng = np.random.default_rng(seed)
    n = n_boxes + n_capsules
    types = np.empty(n, dtype=np.int32)
    types[:n_boxes] = _native.BOX
    types[n_boxes:] = _native.CAPSULE
    R = uniform_rotations(rng, n)
    t = rng.uniform(0.0, extent, (n, 3))
    params = np.zeros((n, 3))
    params[:n_boxes] = np.exp(rng.uniform(np.log(0.01), np.log(0.5), (n_boxes, 3)))
    params[n_boxes:, 0] = rng.uniform(0.05, 0.5, n_capsules)
    params[n_boxes:, 1] = rng.uniform(0.1, 2.0, n_capsules)
    scene = _scene_from(types, R, t, params, np.zeros(n, dtype=np.int64),
                        np.zeros((0, 3), dtype=np.float32))
    return Workload("broad_phase_scene", scene, np.zeros(0, dtype=np.int32),
                    np.zeros(0, dtype=np.int32),
                    {"seed": seed, "n_boxes": n_boxes, "n_capsules": n_capsules})


def serving_workload(seed=4, n_colliders=512, n_requests=50000, rate=20000.0, max_pairs=4096,
                     n_poses=65536, workspace=4.0):
    """(HostScene, poses (P,12), list of serving.Request)."""
    from .serving import Request
    rng = np.random.default_rng(seed)
    kind = np.arange(n_colliders) % 3
    types = np.where(kind == 0, _native.CAPSULE, np.where(kind == 1, _native.BOX, _native.HULL))
    types = types.astype(np.int32)
    R = uniform_rotations(rng, n_colliders)
    t = rng.uniform(0.0, workspace, (n_colliders, 3))
    params = np.zeros((n_colliders, 3))
    cap = types == _native.CAPSULE
    params[cap, 0] = rng.uniform(0.05, 0.3, int(cap.sum()))
    params[cap, 1] = rng.uniform(0.2, 1.0, int(cap.sum()))
    box = types == _native.BOX
    params[box] = rng.uniform(0.1, 1.0, (int(box.sum()), 3))
    hull = types == _native.HULL
    counts = np.zeros(n_colliders, dtype=np.int64)
    counts[hull] = rng.integers(64, 1025, int(hull.sum()))
    total = int(counts.sum())
    owner_scale = np.repeat(rng.uniform(0.2, 0.6, int(hull.sum())), counts[hull])
    v32 = (_sphere_points(rng, total) * (owner_scale * (1.0 + rng.uniform(-0.1, 0.1, total)))[:, None]
           ).astype(np.float32)
    scene = _scene_from(types, R, t, params, counts, v32)
    PR = uniform_rotations(rng, n_poses)
    poses = np.concatenate([PR.reshape(n_poses, 9), rng.uniform(0.0, workspace, (n_poses, 3))],
                           axis=1)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrival = np.cumsum(gaps) - gaps[0]
    sizes = np.clip(np.rint(np.exp(rng.uniform(0.0, np.log(max_pairs), n_requests))), 1,
                    max_pairs).astype(np.int64)
    total_pairs = int(sizes.sum())
    a = rng.integers(0, n_colliders, total_pairs)
    b = (a + rng.integers(1, n_colliders, total_pairs)) % n_colliders
    pa = rng.integers(0, n_poses, total_pairs)
    pb = rng.integers(0, n_poses, total_pairs)
    reqs = []
    off = 0
    for k in range(n_requests):
        e = off + int(sizes[k])
        reqs.append(Request(k, arrival[k], a[off:e], b[off:e], pa[off:e], pb[off:e]))
        off = e
    return scene, poses, reqs


def posed_instances(scene, poses, pair_a, pair_b, pose_a, pose_b):
    """Host-side instance table of posed pairs: rows 2i / 2i+1 are scene rows
    pair_a[i] / pair_b[i] with pose rows pose_a[i] / pose_b[i] (the checker's
    view of d3d_instance_pairs)."""
    n = len(pair_a)
    f = np.empty((2 * n, F64_COLS))
    f[0::2] = scene.shape_f64[pair_a]
    f[1::2] = scene.shape_f64[pair_b]
    P = np.asarray(poses, dtype=np.float64).reshape(-1, 12)
    f[0::2, :12] = P[pose_a]
    f[1::2, :12] = P[pose_b]
    i = np.empty((2 * n, I32_COLS), dtype=np.int32)
    i[0::2] = scene.shape_i32[pair_a]
    i[1::2] = scene.shape_i32[pair_b]
    return HostScene(f, i, scene.vx, scene.vy, scene.vz, scene.v64)
