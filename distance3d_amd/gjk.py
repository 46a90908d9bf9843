"""GJK with the reference's call surface (gjk/__init__.py:17-48,
gjk/_gjk_jolt.py:29-221), evaluated by the gfx950 kernels.

Single-pair calls pack both colliders into a two-collider scene and launch a
batch of one; batched work goes through :func:`gjk_distance_batch` /
:func:`gjk_intersection_batch` on a device-resident scene.
"""
from . import _native, kernels
from .scene import SceneBuilder

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
    simplex = r.simplex.cpu().numpy()[0, 0].copy()
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
           "gjk_intersection_batch", "MAX_FLOAT"]
