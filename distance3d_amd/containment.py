"""Axis-aligned bounding boxes of colliders (containment.py:6-229), computed
by the GPU AABB kernel.  Each function returns ``(mins, maxs)`` like the
reference; ``collider.aabb()`` returns the (3, 2) form."""
import numpy as np

from . import kernels
from .scene import SceneBuilder


def _one(add):
    b = SceneBuilder()
    add(b)
    out = kernels.aabb_batch(b.build().to_device(), [0]).cpu().numpy()[0]
    return out[:, 0], out[:, 1]


def axis_aligned_bounding_box(P):
    return _one(lambda b: b.add_hull(np.asarray(P, dtype=np.float64)))


def sphere_aabb(center, radius):
    return _one(lambda b: b.add_sphere(center, radius))


def box_aabb(box2origin, size):
    return _one(lambda b: b.add_box(box2origin, size))


def cylinder_aabb(cylinder2origin, radius, length):
    return _one(lambda b: b.add_cylinder(cylinder2origin, radius, length))


def capsule_aabb(capsule2origin, radius, height):
    return _one(lambda b: b.add_capsule(capsule2origin, radius, height))


def ellipsoid_aabb(ellipsoid2origin, radii):
    return _one(lambda b: b.add_ellipsoid(ellipsoid2origin, radii))


def disk_aabb(center, radius, normal):
    return _one(lambda b: b.add_disk(center, radius, normal))


def cone_aabb(cone2origin, radius, height):
    return _one(lambda b: b.add_cone(cone2origin, radius, height))


def ellipse_aabb(center, axes, radii):
    return _one(lambda b: b.add_ellipse(center, axes, radii))
