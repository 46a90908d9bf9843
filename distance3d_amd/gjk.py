"""GJK with the reference's call surface (gjk/__init__.py:17-48,
gjk/_gjk_jolt.py:29-221), evaluated by the gfx950 kernels.

Single-pair calls pack both colliders into a two-collider scene and launch a
batch of one; batched work goes through :func:`gjk_distance_batch` /
:func:`gjk_intersection_batch` on a device-resident scene.
"""
import numpy as np

from . import _native, kernels
from .scene import SceneBuilder


class GjkSimplex(np.ndarray):
    """The (4, 3) simplex ``gjk`` returns (Minkowski-difference rows, stale
    rows included as the reference leaves them), carrying what EPA needs to
    seed from it: ``n_points`` valid rows and the support points on both
    colliders (``support1``, ``support2``, each (4, 3))."""

    def __new__(cls, Y, P, Q, n_points):
        obj = np.asarray(Y, dtype=np.float64).reshape(4, 3).copy().view(cls)
        obj.support1 = np.asarray(P, dtype=np.float64).reshape(4, 3).copy()
        obj.support2 = np.asarray(Q, dtype=np.float64).reshape(4, 3).copy()
        obj.n_points = int(n_points)
        return obj

    def __array_finalize__(self, obj):
        if obj is None:
            return
        self.support1 = getattr(obj, "support1", None)
        self.support2 = getattr(obj, "support2", None)
        self.n_points = getattr(obj, "n_points", None)

MAX_FLOAT = _native.MAX_FLOAT

gjk_distance_batch = kernels.gjk_distance_batch
gjk_intersection_batch = kernels.gjk_intersection_batch


def _pair_scene(collider1, collider2):
    b = SceneBuilder()
    i1 = collider1.pack(b)
    i2 = collider2.pack(b)
    return b.build().to_device(), i1, i2


def gjk_distance_jolt(collider1, collider2, tolerance=1e-10, max_distance_squared=100000.0,
                      sanity_check=1e-8):
    """Distance between two convex colliders (gjk/_gjk_jolt.py:138-221).

    Returns ``(distance, closest_point1, closest_point2, simplex)``, or
    ``(MAX_FLOAT, None, None, None)`` when the colliders are further apart
    than ``sqrt(max_distance_squared)``.  ``simplex`` is the (4, 3) array of
    Minkowski-difference support points the loop ended with.
    """
    scene, i1, i2 = _pair_scene(collider1, collider2)
    r = kernels.gjk_distance_batch(scene, [i1], [i2], tolerance, max_distance_squared,
                                   sanity_check, want_simplex=True)
    status = int(r.status.cpu().item())
    if status & 0xFF == _native.CLIPPED:
        return MAX_FLOAT, None, None, None
    if status & _native.BAD_INPUT:
        raise ValueError("invalid collider")
    if status & _native.SANITY_FAILED:
        raise AssertionError("Sanity check failed")
    if status & 0xFF == _native.ITER_CAP:
        raise RuntimeError("GJK did not terminate within the iteration cap")
    dist = float(r.dist.cpu().item())
    cp1 = r.cp_a.cpu().numpy()[0]
    cp2 = r.cp_b.cpu().numpy()[0]
    S = r.simplex.cpu().numpy()[0]
    simplex = GjkSimplex(S[0], S[1], S[2], int(r.n_points.cpu().item()))
    return dist, cp1, cp2, simplex


def gjk_intersection_jolt(collider1, collider2, tolerance=1e-10):
    """Do two convex colliders intersect? (gjk/_gjk_jolt.py:29-80)"""
    scene, i1, i2 = _pair_scene(collider1, collider2)
    hit, status = kernels.gjk_intersection_batch(scene, [i1], [i2], tolerance, want_status=True)
    st = int(status.cpu().item())
    if st & _native.BAD_INPUT:
        raise ValueError("invalid collider")
    if st == _native.ITER_CAP:
        raise RuntimeError("GJK did not terminate within the iteration cap")
    return bool(hit.cpu().item())


gjk = gjk_distance = gjk_distance_jolt
gjk_intersection = gjk_intersection_jolt


def gjk_distance_jolt_iterations(collider1, collider2, tolerance=1e-10,
                                 max_distance_squared=100000.0):
    """Number of GJK iterations (gjk/_gjk_jolt.py:714-785)."""
    scene, i1, i2 = _pair_scene(collider1, collider2)
    r = kernels.gjk_distance_batch(scene, [i1], [i2], tolerance, max_distance_squared,
                                   float("inf"))
    return int(r.iterations.cpu().item())


__all__ = ["gjk", "gjk_distance", "gjk_intersection", "gjk_distance_jolt",
           "gjk_intersection_jolt", "gjk_distance_jolt_iterations", "gjk_distance_batch",
           "gjk_intersection_batch", "MAX_FLOAT", "GjkSimplex"]
