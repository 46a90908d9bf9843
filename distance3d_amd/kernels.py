"""Batched entry points over a device-resident scene (torch tensors in, torch
tensors out).  Each function is one call into libd3d_hip on the current
stream; none of them synchronises."""
import numpy as np

from . import _native
from ._native import check, lib, ptr, stream_ptr
from .scene import scene_descriptor


def _dev_i32(x, device):
    import torch
    return torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x,
                           dtype=torch.int32).to(device).contiguous()


def _dev_f64(x, device):
    import torch
    return torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x,
                           dtype=torch.float64).to(device).contiguous()


def support_batch(scene, shape_ids, dirs, stream=None):
    """Support points of ``scene`` colliders ``shape_ids`` along ``dirs``
    (n, 3) -> (n, 3) float64 tensor."""
    import torch
    ids = _dev_i32(shape_ids, scene.device)
    d = _dev_f64(dirs, scene.device).reshape(-1, 3)
    n = ids.numel()
    if d.shape[0] != n:
        raise ValueError("shape_ids and dirs differ in length")
    out = torch.empty((n, 3), dtype=torch.float64, device=scene.device)
    check(lib().d3d_support_batch(scene.c, ptr(ids), ptr(d), n, ptr(out), stream_ptr(stream)),
          "d3d_support_batch")
    return out


def aabb_batch(scene, shape_ids, stream=None):
    """AABBs (n, 3, 2) of ``scene`` colliders ``shape_ids``."""
    import torch
    ids = _dev_i32(shape_ids, scene.device)
    n = ids.numel()
    out = torch.empty((n, 3, 2), dtype=torch.float64, device=scene.device)
    check(lib().d3d_aabb_batch(scene.c, ptr(ids), n, ptr(out), stream_ptr(stream)),
          "d3d_aabb_batch")
    return out


class GjkResult:
    """Outputs of :func:`gjk_distance_batch` (device tensors)."""

    def __init__(self, dist, cp_a, cp_b, status, iterations, simplex, n_points):
        self.dist = dist
        self.cp_a = cp_a
        self.cp_b = cp_b
        self.status = status
        self.iterations = iterations
        self.simplex = simplex
        self.n_points = n_points

    @property
    def intersecting(self):
        """dist == 0.0, the reference's intersection criterion."""
        return self.dist == 0.0


def gjk_distance_batch(scene, pair_a, pair_b, tolerance=1e-10, max_distance_squared=100000.0,
                       sanity_check=1e-8, max_iterations=128, lanes_per_pair=None,
                       want_simplex=False, want_iterations=True, out=None, workspace=None,
                       stream=None, launch=None):
    """GJK distance over pairs (gjk.gjk semantics per pair).

    ``out`` may hold preallocated tensors (dist, cp_a, cp_b, status,
    iterations, simplex, n_points) and ``workspace`` a uint8 device tensor of
    at least :func:`gjk_workspace_bytes` bytes, so a loop allocates nothing.
    ``lanes_per_pair`` (8 entries, one per work class) and ``launch`` (dict:
    ``blocks_per_cu``, ``claim_factor``) override the launch tuning."""
    import torch
    a = _dev_i32(pair_a, scene.device)
    b = _dev_i32(pair_b, scene.device)
    n = a.numel()
    if b.numel() != n:
        raise ValueError("pair_a and pair_b differ in length")
    dev = scene.device
    if out is None:
        out = GjkResult(
            torch.empty(n, dtype=torch.float64, device=dev),
            torch.empty((n, 3), dtype=torch.float64, device=dev),
            torch.empty((n, 3), dtype=torch.float64, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev),
            torch.empty(n, dtype=torch.int32, device=dev) if want_iterations else None,
            torch.empty((n, 3, 4, 3), dtype=torch.float64, device=dev) if want_simplex else None,
            torch.empty(n, dtype=torch.int32, device=dev) if want_simplex else None)
    prm = _native.gjk_params(tolerance, max_distance_squared, sanity_check, max_iterations,
                             lanes_per_pair=lanes_per_pair, launch=launch)
    import ctypes
    if workspace is None:
        check(lib().d3d_gjk_distance_batch(
            scene.c, ptr(a), ptr(b), n, ctypes.byref(prm), ptr(out.dist), ptr(out.cp_a),
            ptr(out.cp_b), ptr(out.status), ptr(out.iterations), ptr(out.simplex),
            ptr(out.n_points), stream_ptr(stream)), "d3d_gjk_distance_batch")
    else:
        check(lib().d3d_gjk_distance_batch_ws(
            scene.c, ptr(a), ptr(b), n, ctypes.byref(prm), ptr(out.dist), ptr(out.cp_a),
            ptr(out.cp_b), ptr(out.status), ptr(out.iterations), ptr(out.simplex),
            ptr(out.n_points), ptr(workspace), workspace.numel(), stream_ptr(stream)),
            "d3d_gjk_distance_batch_ws")
    return out


def gjk_workspace_bytes(n):
    """Device workspace bytes the batched GJK needs for n pairs."""
    return int(_native.load_library().d3d_gjk_workspace_bytes(int(n)))


def gjk_intersection_batch(scene, pair_a, pair_b, tolerance=1e-10, max_iterations=128,
                           want_status=False, stream=None):
    """GJK intersection test over pairs -> (hit uint8 tensor, status or None)."""
    import torch
    a = _dev_i32(pair_a, scene.device)
    b = _dev_i32(pair_b, scene.device)
    n = a.numel()
    if b.numel() != n:
        raise ValueError("pair_a and pair_b differ in length")
    hit = torch.empty(n, dtype=torch.uint8, device=scene.device)
    status = torch.empty(n, dtype=torch.int32, device=scene.device) if want_status else None
    check(lib().d3d_gjk_intersection_batch(
        scene.c, ptr(a), ptr(b), n, float(tolerance), int(max_iterations), ptr(hit),
        ptr(status), stream_ptr(stream)), "d3d_gjk_intersection_batch")
    return hit, status


def mpr_batch(scene, pair_a, pair_b, centers, tolerance=0.0001, max_iterations=100,
              penetration=True, stream=None):
    """MPR over pairs (d3d_mpr_batch).  ``centers``: (n_shapes, 3) collider
    centres.  Returns (status int32 tensor: 1 intersecting, 0 separated,
    0x100 refinement cap, 0x200 bad record; out (n, 7) float64 tensor:
    depth, direction, contact position, NaN when separated; None when
    ``penetration`` is false)."""
    import torch
    a = _dev_i32(pair_a, scene.device)
    b = _dev_i32(pair_b, scene.device)
    n = a.numel()
    if b.numel() != n:
        raise ValueError("pair_a and pair_b differ in length")
    c = torch.as_tensor(centers, dtype=torch.float64).reshape(-1, 3).to(scene.device).contiguous()
    if c.shape[0] != scene.n_shapes:
        raise ValueError("centers must hold one row per collider (%d), got %d"
                         % (scene.n_shapes, c.shape[0]))
    status = torch.empty(n, dtype=torch.int32, device=scene.device)
    out = torch.empty((n, 7), dtype=torch.float64, device=scene.device) if penetration else None
    check(lib().d3d_mpr_batch(scene.c, ptr(a), ptr(b), ptr(c), n, float(tolerance),
                              int(max_iterations), 1 if penetration else 0, ptr(status), ptr(out),
                              stream_ptr(stream)), "d3d_mpr_batch")
    return status, out


TET_MAX_POLY = 24


def tet_pairs_batch(T1, e1, X1, T2, e2, X2, pairs, youngs_modulus1=1.0, youngs_modulus2=1.0,
                    stream=None, geometry=True, checked=True):
    """Contact plane and polygon of tetrahedron pairs (d3d_tet_pairs_batch).
    T (n,4,3), e (n,4), X (n,4,4) per body, pairs (m,2); tensors or arrays.
    Returns device tensors (status (m,), planes (m,4), n_vertices (m,),
    polygons (m,24,3)); planes and polygons are None when not ``geometry``.
    ``checked=False`` skips the pair index range check (a device reduction
    and a host read) for pairs produced in range by this library."""
    import torch

    def dev(a, shape):
        t = torch.as_tensor(a, dtype=torch.float64)
        return (t if t.is_cuda else t.cuda()).reshape(shape).contiguous()

    T1, T2 = dev(T1, (-1, 4, 3)), dev(T2, (-1, 4, 3))
    e1, e2 = dev(e1, (-1, 4)), dev(e2, (-1, 4))
    X1, X2 = dev(X1, (-1, 4, 4)), dev(X2, (-1, 4, 4))
    pr = torch.as_tensor(pairs, dtype=torch.int32)
    pr = (pr if pr.is_cuda else pr.cuda()).reshape(-1, 2).contiguous()
    m = pr.shape[0]
    if checked and m and (int(pr[:, 0].max()) >= T1.shape[0] or int(pr[:, 1].max()) >= T2.shape[0]
                          or int(pr.min()) < 0):
        raise ValueError("pair index out of range")
    status = torch.empty(m, dtype=torch.int32, device="cuda")
    planes = torch.empty((m, 4), dtype=torch.float64, device="cuda") if geometry else None
    nv = torch.empty(m, dtype=torch.int32, device="cuda")
    polys = torch.empty((m, TET_MAX_POLY, 3), dtype=torch.float64, device="cuda") if geometry else None
    ws = torch.empty(max(int(_native.load_library().d3d_tet_pairs_workspace_bytes(m)), 1),
                     dtype=torch.uint8, device="cuda")
    check(lib().d3d_tet_pairs_batch(ptr(T1), ptr(e1), ptr(X1), ptr(T2), ptr(e2), ptr(X2), ptr(pr),
                                    m, float(youngs_modulus1), float(youngs_modulus2), ptr(status),
                                    ptr(planes), ptr(nv), ptr(polys), ptr(ws), ws.numel(),
                                    stream_ptr(stream)), "d3d_tet_pairs_batch")
    return status, planes, nv, polys


def _dev64(a, shape):
    import torch
    t = torch.as_tensor(a, dtype=torch.float64)
    return (t if t.is_cuda else t.cuda()).reshape(shape).contiguous()


def _dev32(a, shape):
    import torch
    t = torch.as_tensor(a, dtype=torch.int32)
    return (t if t.is_cuda else t.cuda()).reshape(shape).contiguous()


def polygon_angles(y, x, stream=None):
    """atan2(y, x) rounded correctly, elementwise (d3d_polygon_angles): the
    angle the contact-polygon ordering uses.  Returns a device tensor."""
    import torch
    y = _dev64(y, (-1,))
    x = _dev64(x, (-1,))
    if y.shape != x.shape:
        raise ValueError("y and x differ in length")
    out = torch.empty_like(y)
    check(lib().d3d_polygon_angles(ptr(y), ptr(x), y.shape[0], ptr(out), stream_ptr(stream)),
          "d3d_polygon_angles")
    return out


def tet_mesh_prep(T, aabbs=True, transforms=True, stream=None):
    """AABBs (n,3,2) and barycentric transforms (n,4,4) of tetrahedra T
    (n,4,3) (d3d_tet_mesh_prep).  Returns device tensors (aabbs or None, X or
    None, flags (n,) int32: 1 for a degenerate tetrahedron)."""
    import torch
    T = _dev64(T, (-1, 4, 3))
    n = T.shape[0]
    ab = torch.empty((n, 3, 2), dtype=torch.float64, device="cuda") if aabbs else None
    X = torch.empty((n, 4, 4), dtype=torch.float64, device="cuda") if transforms else None
    flags = torch.empty(n, dtype=torch.int32, device="cuda")
    check(lib().d3d_tet_mesh_prep(ptr(T), n, ptr(ab) if ab is not None else None,
                                  ptr(X) if X is not None else None, ptr(flags),
                                  stream_ptr(stream)), "d3d_tet_mesh_prep")
    return ab, X, flags


def tet_contact_forces(T1, e1, pairs, status, planes, n_vertices, polygons, youngs_modulus1=1.0,
                       stream=None, checked=True):
    """Centre, force and area of each intersecting pair's contact polygon
    (d3d_tet_contact_forces); inputs as tet_pairs_batch takes and returns
    them.  Returns a device tensor (m, 7): centre (3), force (3), area.
    ``checked=False`` skips the index and vertex-count range checks for
    inputs produced by tet_pairs_batch on in-range pairs."""
    import torch
    T1 = _dev64(T1, (-1, 4, 3))
    e1 = _dev64(e1, (-1, 4))
    pr = _dev32(pairs, (-1, 2))
    m = pr.shape[0]
    st = _dev32(status, (m,))
    pl = _dev64(planes, (m, 4))
    nv = _dev32(n_vertices, (m,))
    po = _dev64(polygons, (m, TET_MAX_POLY, 3))
    if checked and m and (int(pr[:, 0].max()) >= T1.shape[0] or int(pr[:, 0].min()) < 0):
        raise ValueError("pair index out of range")
    if checked and m and int(nv.max()) > TET_MAX_POLY:
        raise ValueError("polygon vertex count out of range")
    out = torch.empty((m, 7), dtype=torch.float64, device="cuda")
    check(lib().d3d_tet_contact_forces(ptr(T1), ptr(e1), ptr(pr), ptr(st), ptr(pl), ptr(nv), ptr(po),
                                       m, float(youngs_modulus1), ptr(out), stream_ptr(stream)),
          "d3d_tet_contact_forces")
    return out


class EpaResult:
    """Outputs of :func:`epa_batch` (device tensors); ``faces`` (n,
    max_faces, 4, 3) and ``n_faces`` (n) only with ``want_faces``."""

    def __init__(self, out, status, iterations, faces=None, n_faces=None):
        self.faces = faces
        self.n_faces = n_faces
        self.raw = out
        self.depth = out[:, 0]
        self.normal = out[:, 1:4]
        self.mtv = out[:, 4:7]
        self.cp_a = out[:, 7:10]
        self.cp_b = out[:, 10:13]
        self.status = status
        self.iterations = iterations

    @property
    def converged(self):
        return self.status == 1


def epa_batch(scene, pair_a, pair_b, simplex, n_points, max_iterations=64, max_faces=128,
              epsilon=1e-8, stream=None, want_faces=False):
    """EPA over pairs seeded from GJK simplices ((n,3,4,3) Y/P/Q rows and
    n_points from :func:`gjk_distance_batch` with ``want_simplex=True``).
    With ``want_faces`` the final polytope of every pair is returned too."""
    import torch
    a = _dev_i32(pair_a, scene.device)
    b = _dev_i32(pair_b, scene.device)
    n = a.numel()
    if b.numel() != n:
        raise ValueError("pair_a and pair_b differ in length")
    s = _dev_f64(simplex, scene.device).reshape(n, 36)
    npts = _dev_i32(n_points, scene.device)
    out = torch.empty((n, 13), dtype=torch.float64, device=scene.device)
    status = torch.empty(n, dtype=torch.int32, device=scene.device)
    iters = torch.empty(n, dtype=torch.int32, device=scene.device)
    faces = n_faces = None
    if want_faces:
        faces = torch.zeros((n, max(int(max_faces), 0), 4, 3), dtype=torch.float64,
                            device=scene.device)
        n_faces = torch.zeros(n, dtype=torch.int32, device=scene.device)
    check(lib().d3d_epa_batch(scene.c, ptr(a), ptr(b), n, ptr(s), ptr(npts),
                              int(max_iterations), int(max_faces), float(epsilon), ptr(out),
                              ptr(status), ptr(iters), ptr(faces), ptr(n_faces),
                              stream_ptr(stream)), "d3d_epa_batch")
    return EpaResult(out, status, iters, faces, n_faces)


class Bvh:
    """A linear BVH over ``aabbs`` (n,3,2) float64 device tensor, held in a
    device buffer (uint8 tensor).  The boxes must stay unchanged while the
    BVH is queried."""

    def __init__(self, aabbs, buffer, n):
        self.aabbs = aabbs
        self.buffer = buffer
        self.n = n


def bvh_build(aabbs, stream=None):
    """Build a BVH over (n,3,2) float64 boxes (tensor or array)."""
    import torch
    a = torch.as_tensor(aabbs, dtype=torch.float64)
    a = a.cuda() if not a.is_cuda else a
    a = a.reshape(-1, 3, 2).contiguous()
    n = a.shape[0]
    if n < 1:
        raise ValueError("a BVH needs at least one box")
    buf = torch.empty(int(_native.load_library().d3d_bvh_bytes(n)), dtype=torch.uint8,
                      device=a.device)
    check(lib().d3d_bvh_build(ptr(a), n, ptr(buf), buf.numel(), stream_ptr(stream)),
          "d3d_bvh_build")
    return Bvh(a, buf, n)


def bvh_refit(bvh, aabbs, stream=None):
    """New (n,3,2) float64 boxes for the leaves of ``bvh`` (box i keeps leaf
    slot i, the topology stays); the boxes are copied into ``bvh.aabbs``."""
    import torch
    a = torch.as_tensor(aabbs, dtype=torch.float64).reshape(-1, 3, 2)
    if a.shape[0] != bvh.n:
        raise ValueError("refit needs %d boxes, got %d" % (bvh.n, a.shape[0]))
    bvh.aabbs.copy_(a.to(bvh.aabbs.device))
    check(lib().d3d_bvh_refit(ptr(bvh.aabbs), bvh.n, ptr(bvh.buffer), bvh.buffer.numel(),
                              stream_ptr(stream)), "d3d_bvh_refit")
    return bvh


# below this many box tests (n * m) overlap_dense beats a BVH query
DENSE_OVERLAP_MAX_TESTS = 1 << 28


def overlap_dense(a, b, cap=None, counter=None, stream=None):
    """All (i, j) with closed-interval overlap of a[i] and b[j], by testing
    every pair (d3d_overlap_dense), row-major (i, then j ascending), as an
    (k, 2) int32 device tensor.  Reads the pair count back (one host
    synchronisation); a second pass runs when ``cap`` was too small."""
    import torch
    a = _dev64(a, (-1, 3, 2))
    b = _dev64(b, (-1, 3, 2)).to(a.device)
    n, m = a.shape[0], b.shape[0]
    if counter is None:
        counter = torch.zeros(1, dtype=torch.int64, device=a.device)
    ws = torch.empty(max(int(_native.load_library().d3d_overlap_dense_workspace_bytes(n)), 8),
                     dtype=torch.uint8, device=a.device)
    if cap is None:
        cap = 4 * n + 1024
    while True:
        pairs = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=a.device)
        check(lib().d3d_overlap_dense(ptr(a), n, ptr(b), m, ptr(pairs), cap, ptr(counter), ptr(ws),
                                      ws.numel(), stream_ptr(stream)), "d3d_overlap_dense")
        k = int(counter.item())
        if k <= cap:
            return pairs[:k]
        cap = k


def bvh_query(bvh, queries, cap=None, counter=None, stream=None, sort=True):
    """All (query, box) index pairs with closed-interval overlap, as an
    (k, 2) int32 device tensor (lexicographically sorted when ``sort``).

    Reads the pair count back to size the output (one host synchronisation);
    a second launch runs when ``cap`` was too small."""
    import torch
    q = torch.as_tensor(queries, dtype=torch.float64)
    q = q.to(bvh.aabbs.device).reshape(-1, 3, 2).contiguous()
    m = q.shape[0]
    if counter is None:
        counter = torch.zeros(1, dtype=torch.int64, device=q.device)
    if cap is None:
        cap = 4 * m + 1024
    while True:
        pairs = torch.empty((max(cap, 1), 2), dtype=torch.int32, device=q.device)
        check(lib().d3d_bvh_query(ptr(bvh.buffer), bvh.n, ptr(bvh.aabbs), ptr(q), m, ptr(pairs),
                                  cap, ptr(counter), stream_ptr(stream)), "d3d_bvh_query")
        k = int(counter.item())
        if k <= cap:
            break
        cap = k
    pairs = pairs[:k]
    if sort and k > 1:
        key = pairs[:, 0].to(torch.int64) * bvh.n + pairs[:, 1].to(torch.int64)
        pairs = pairs[torch.argsort(key)]
    return pairs


class SceneView:
    """A collider table (device tensors) over another scene's vertex pools;
    used for per-pair pose instances.  Duck-types DeviceScene for the batch
    functions above."""

    def __init__(self, base, shape_f64, shape_i32, n_shapes):
        self.device = base.device
        self.base = base
        self.shape_f64 = shape_f64
        self.shape_i32 = shape_i32
        self.n_shapes = n_shapes
        self._c = scene_descriptor(base, shape_f64, shape_i32, n_shapes)

    @property
    def c(self):
        import ctypes
        return ctypes.byref(self._c)


def instance_pairs(scene, pair_a, pair_b, pose_a=None, pose_b=None, pose_idx_a=None,
                   pose_idx_b=None, n_poses=0, out=None, stream=None):
    """Posed pair instances of ``scene`` colliders (see d3d_instance_pairs).
    Returns (SceneView over 2n instances, pair_a', pair_b') with the pair
    lists 2i / 2i+1.  ``out`` may pass preallocated (f64 (2n,16), i32 (2n,4),
    a (n), b (n)) tensors."""
    import torch
    n = int(pair_a.numel())
    dev = scene.device
    if out is None:
        out = (torch.empty((2 * n, 16), dtype=torch.float64, device=dev),
               torch.empty((2 * n, 4), dtype=torch.int32, device=dev),
               torch.empty(n, dtype=torch.int32, device=dev),
               torch.empty(n, dtype=torch.int32, device=dev))
    f64, i32, ia, ib = out
    check(lib().d3d_instance_pairs(scene.c, ptr(pair_a), ptr(pair_b), ptr(pose_a), ptr(pose_b),
                                   ptr(pose_idx_a), ptr(pose_idx_b), int(n_poses), n, ptr(f64),
                                   ptr(i32), ptr(ia), ptr(ib), stream_ptr(stream)),
          "d3d_instance_pairs")
    return SceneView(scene, f64, i32, 2 * n), ia[:n], ib[:n]


def gjk_epa_workspace_bytes(n):
    return int(_native.load_library().d3d_gjk_epa_workspace_bytes(int(n)))


class GjkEpaResult:
    """Outputs of :func:`gjk_epa_batch` (device tensors)."""

    def __init__(self, dist, cp_a, cp_b, status, epa_out, epa_status):
        self.dist = dist
        self.cp_a = cp_a
        self.cp_b = cp_b
        self.status = status
        self.epa_out = epa_out
        self.epa_status = epa_status

    @staticmethod
    def empty(n, device):
        import torch
        return GjkEpaResult(torch.empty(n, dtype=torch.float64, device=device),
                            torch.empty((n, 3), dtype=torch.float64, device=device),
                            torch.empty((n, 3), dtype=torch.float64, device=device),
                            torch.empty(n, dtype=torch.int32, device=device),
                            torch.empty((n, 13), dtype=torch.float64, device=device),
                            torch.empty(n, dtype=torch.int32, device=device))


def gjk_epa_batch(scene, pair_a, pair_b, out=None, workspace=None, lanes_per_pair=None,
                  epa_max_iterations=64, epa_max_faces=128, epa_epsilon=1e-8, stream=None,
                  iterations=None, **gjk_kw):
    """GJK distance for every pair and EPA for the intersecting ones, in one
    device pipeline (no host synchronisation).  ``epa_status`` is -1 where
    GJK found no intersection.  ``iterations`` (optional int32 device tensor
    of n entries) receives each pair's GJK iteration count."""
    import ctypes
    import torch
    a = _dev_i32(pair_a, scene.device)
    b = _dev_i32(pair_b, scene.device)
    n = a.numel()
    if out is None:
        out = GjkEpaResult.empty(n, scene.device)
    if workspace is None:
        workspace = torch.empty(gjk_epa_workspace_bytes(n), dtype=torch.uint8, device=scene.device)
    prm = _native.gjk_params(lanes_per_pair=lanes_per_pair, **gjk_kw)
    check(lib().d3d_gjk_epa_batch_ws(
        scene.c, ptr(a), ptr(b), n, ctypes.byref(prm), int(epa_max_iterations),
        int(epa_max_faces), float(epa_epsilon), ptr(out.dist), ptr(out.cp_a), ptr(out.cp_b),
        ptr(out.status), ptr(iterations), ptr(out.epa_out), ptr(out.epa_status), ptr(workspace),
        workspace.numel(), stream_ptr(stream)), "d3d_gjk_epa_batch_ws")
    return out
