"""ctypes wrapper of the FP64 CPU oracle (d3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libd3d_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_num_threads.restype = ctypes.c_int
        _lib.orc_all_aabbs_overlap.restype = ctypes.c_int64
        _lib.orc_overlap_sorted.restype = ctypes.c_int64
        _lib.orc_intersect_halfplanes.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def set_threads(n):
    lib().orc_set_threads(ctypes.c_int(int(n)))


def num_threads():
    return lib().orc_num_threads()


def _scene_args(scene):
    f, i, vx, vy, vz, v64 = scene.oracle_args()
    return [_p(f), _p(i), _p(vx), _p(vy), _p(vz), _p(v64)]


def support(scene, shape_ids, dirs):
    ids = np.ascontiguousarray(shape_ids, dtype=np.int32)
    d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    out = np.empty((len(ids), 3))
    lib().orc_support(*_scene_args(scene), _p(ids), _p(d), ctypes.c_int64(len(ids)), _p(out))
    return out


def aabbs(scene, shape_ids):
    ids = np.ascontiguousarray(shape_ids, dtype=np.int32)
    out = np.empty((len(ids), 3, 2))
    lib().orc_aabbs(*_scene_args(scene), _p(ids), ctypes.c_int64(len(ids)), _p(out))
    return out


def gjk_distance(scene, pair_a, pair_b, tolerance=1e-10, max_distance_squared=100000.0,
                 sanity_check=1e-8, max_iterations=128, want_simplex=False):
    a = np.ascontiguousarray(pair_a, dtype=np.int32)
    b = np.ascontiguousarray(pair_b, dtype=np.int32)
    n = len(a)
    dist = np.empty(n)
    cpa = np.empty((n, 3))
    cpb = np.empty((n, 3))
    status = np.empty(n, dtype=np.int32)
    iters = np.empty(n, dtype=np.int32)
    simplex = np.empty((n, 3, 4, 3)) if want_simplex else None
    npts = np.empty(n, dtype=np.int32) if want_simplex else None
    lib().orc_gjk_distance(*_scene_args(scene), _p(a), _p(b), ctypes.c_int64(n),
                           ctypes.c_double(tolerance), ctypes.c_double(max_distance_squared),
                           ctypes.c_double(sanity_check), ctypes.c_int(max_iterations),
                           _p(dist), _p(cpa), _p(cpb), _p(status), _p(iters), _p(simplex),
                           _p(npts))
    out = {"dist": dist, "cp_a": cpa, "cp_b": cpb, "status": status, "iterations": iters}
    if want_simplex:
        out["simplex"] = simplex
        out["n_points"] = npts
    return out


def gjk_intersection(scene, pair_a, pair_b, tolerance=1e-10, max_iterations=128):
    a = np.ascontiguousarray(pair_a, dtype=np.int32)
    b = np.ascontiguousarray(pair_b, dtype=np.int32)
    n = len(a)
    hit = np.empty(n, dtype=np.int32)
    iters = np.empty(n, dtype=np.int32)
    lib().orc_gjk_intersection(*_scene_args(scene), _p(a), _p(b), ctypes.c_int64(n),
                               ctypes.c_double(tolerance), ctypes.c_int(max_iterations),
                               _p(hit), _p(iters))
    return hit, iters


def epa(scene, pair_a, pair_b, simplex, n_points, max_iterations=64, max_faces=128,
        epsilon=1e-8, want_faces=False):
    """EPA per pair; with ``want_faces`` also "faces" (n, max_faces, 4, 3),
    the final polytope (rows past "n_faces" are zero)."""
    a = np.ascontiguousarray(pair_a, dtype=np.int32)
    b = np.ascontiguousarray(pair_b, dtype=np.int32)
    n = len(a)
    s = np.ascontiguousarray(simplex, dtype=np.float64).reshape(n, 36)
    npts = np.ascontiguousarray(n_points, dtype=np.int32)
    out = np.empty((n, 13))
    status = np.empty(n, dtype=np.int32)
    iters = np.empty(n, dtype=np.int32)
    faces = np.zeros((n, max_faces, 4, 3)) if want_faces else None
    n_faces = np.zeros(n, dtype=np.int32)
    lib().orc_epa(*_scene_args(scene), _p(a), _p(b), ctypes.c_int64(n), _p(s), _p(npts),
                  ctypes.c_int(max_iterations), ctypes.c_int(max_faces),
                  ctypes.c_double(epsilon), _p(out), _p(status), _p(iters), _p(faces),
                  _p(n_faces))
    r = {"depth": out[:, 0], "normal": out[:, 1:4], "mtv": out[:, 4:7],
         "cp_a": out[:, 7:10], "cp_b": out[:, 10:13], "status": status,
         "iterations": iters, "n_faces": n_faces}
    if want_faces:
        r["faces"] = faces
    return r


def mpr(scene, pair_a, pair_b, centers, tolerance=0.0001, max_iterations=100, penetration=True):
    """MPR per pair (mpr.py:20-109).  ``centers``: (n_shapes, 3) collider
    centres.  Returns (status, out): status 1 intersecting, 0 separated,
    0x100 refinement cap; out (n, 7) depth, direction, contact position
    (NaN when separated)."""
    a = np.ascontiguousarray(pair_a, dtype=np.int32)
    b = np.ascontiguousarray(pair_b, dtype=np.int32)
    c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
    n = len(a)
    status = np.empty(n, dtype=np.int32)
    out = np.empty((n, 7))
    lib().orc_mpr(*_scene_args(scene), _p(a), _p(b), _p(c), ctypes.c_int64(n),
                  ctypes.c_double(tolerance), ctypes.c_int(max_iterations),
                  ctypes.c_int(1 if penetration else 0), _p(status), _p(out))
    return status, out


TET_MAX_POLY = 24


def tet_pairs(T1, e1, X1, T2, e2, X2, pairs, youngs_modulus1=1.0, youngs_modulus2=1.0):
    """Contact plane and polygon per tetrahedron pair
    (_tetrahedron_intersection.py:89-139).  Returns (status (m,) int32, bit 0
    intersecting, 0x10 same-tetrahedron case, 0x20 the plane cuts both,
    0x100 half-plane overflow;
    planes (m, 4); n_vertices (m,); polygons (m, 24, 3), NaN past n)."""
    c = lambda a, s: np.ascontiguousarray(a, dtype=np.float64).reshape(s)
    T1 = c(T1, (-1, 4, 3)); T2 = c(T2, (-1, 4, 3))
    e1 = c(e1, (-1, 4)); e2 = c(e2, (-1, 4))
    X1 = c(X1, (-1, 4, 4)); X2 = c(X2, (-1, 4, 4))
    pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    m = len(pr)
    status = np.empty(m, dtype=np.int32)
    planes = np.empty((m, 4))
    nv = np.empty(m, dtype=np.int32)
    polys = np.empty((m, TET_MAX_POLY, 3))
    lib().orc_tet_pairs(_p(T1), _p(e1), _p(X1), _p(T2), _p(e2), _p(X2), _p(pr), ctypes.c_int64(m),
                        ctypes.c_double(youngs_modulus1), ctypes.c_double(youngs_modulus2),
                        _p(status), _p(planes), _p(nv), _p(polys))
    return status, planes, nv, polys


def atan2(y, x):
    """The oracle's correctly rounded atan2 (the polygon ordering's angle)."""
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty(len(y))
    lib().orc_atan2(_p(y), _p(x), ctypes.c_int64(len(y)), _p(out))
    return out


def tet_contact_forces(T1, e1, pairs, status, planes, nv, polys, youngs_modulus1=1.0):
    """compute_contact_force (_forces.py:52-111) for every pair with status
    bit 0, from the first body's tetrahedra: (m, 7) rows of centre (3),
    force (3), area; zero rows for the other pairs."""
    c = lambda a, s, t=np.float64: np.ascontiguousarray(a, dtype=t).reshape(s)
    T1 = c(T1, (-1, 4, 3)); e1 = c(e1, (-1, 4))
    pr = c(pairs, (-1, 2), np.int32)
    m = len(pr)
    st = c(status, (m,), np.int32); pl = c(planes, (m, 4)); n = c(nv, (m,), np.int32)
    po = c(polys, (m, TET_MAX_POLY, 3))
    out = np.empty((m, 7))
    lib().orc_tet_contact_forces(_p(T1), _p(e1), _p(pr), _p(st), _p(pl), _p(n), _p(po),
                                 ctypes.c_int64(m), ctypes.c_double(youngs_modulus1), _p(out))
    return out


def tet_mesh_prep(T):
    """Per tetrahedron: AABB (n, 3, 2) (_mesh_processing.py:4-20), barycentric
    transform (n, 4, 4) by a partial-pivot LU inverse, degenerate flag (n,)."""
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1, 4, 3)
    n = len(T)
    aabbs = np.empty((n, 3, 2)); X = np.empty((n, 4, 4)); flags = np.empty(n, dtype=np.int32)
    lib().orc_tet_mesh_prep(_p(T), ctypes.c_int64(n), _p(aabbs), _p(X), _p(flags))
    return aabbs, X, flags


def intersect_halfplanes(halfplanes):
    """_halfplanes.py:29-64; raises AssertionError where the reference does."""
    hp = np.ascontiguousarray(halfplanes, dtype=np.float64).reshape(-1, 4)
    pts = np.empty((max(3 * len(hp), 1), 2))
    n = lib().orc_intersect_halfplanes(_p(hp), ctypes.c_int(len(hp)), _p(pts))
    if n < 0:
        raise AssertionError("more half-plane intersections than the reference reserves")
    return pts[:n]


def closest_point_simplex(points):
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.empty(3)
    s = ctypes.c_int32()
    lib().orc_closest_point_simplex(_p(p), ctypes.c_int(len(p)), _p(out), ctypes.byref(s))
    return out, s.value


def barycentric(points):
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(p))
    lib().orc_barycentric(_p(p), ctypes.c_int(len(p)), _p(out))
    return out


def all_aabbs_overlap(aabbs1, aabbs2, cap=None):
    """Brute-force closed-interval pairs (i, j), row-major (i outer)."""
    a = np.ascontiguousarray(aabbs1, dtype=np.float64).reshape(-1, 3, 2)
    b = np.ascontiguousarray(aabbs2, dtype=np.float64).reshape(-1, 3, 2)
    if cap is None:
        cap = len(a) * len(b)
    pairs = np.empty((max(cap, 1), 2), dtype=np.int64)
    cnt = lib().orc_all_aabbs_overlap(_p(a), ctypes.c_int64(len(a)), _p(b),
                                      ctypes.c_int64(len(b)), _p(pairs), ctypes.c_int64(cap))
    return pairs[:min(cnt, cap)], cnt


def overlap_pairs_sorted(aabbs1, aabbs2, cap):
    """Same set as all_aabbs_overlap for large inputs (x-sorted sweep)."""
    a = np.ascontiguousarray(aabbs1, dtype=np.float64).reshape(-1, 3, 2)
    b = np.ascontiguousarray(aabbs2, dtype=np.float64).reshape(-1, 3, 2)
    order = np.argsort(b[:, 0, 0], kind="stable").astype(np.int64)
    xmin = np.ascontiguousarray(b[order, 0, 0])
    width = float(np.max(b[:, 0, 1] - b[:, 0, 0])) * (1.0 + 1e-9) + 1e-300
    pairs = np.empty((max(cap, 1), 2), dtype=np.int64)
    cnt = lib().orc_overlap_sorted(_p(a), ctypes.c_int64(len(a)), _p(b), _p(order), _p(xmin),
                                   ctypes.c_int64(len(b)), ctypes.c_double(width), _p(pairs),
                                   ctypes.c_int64(cap))
    return pairs[:min(cnt, cap)], cnt
