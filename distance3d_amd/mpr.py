"""Minkowski portal refinement with the reference's call surface
(mpr.py:20-109), evaluated by the gfx950 kernel (d3d_mpr_batch).

Single-pair calls pack both colliders into a two-collider scene and launch a
batch of one; batched work goes through :func:`mpr_batch` on a
device-resident scene with one centre per collider.
"""
import numpy as np

from . import kernels
from .scene import SceneBuilder

mpr_batch = kernels.mpr_batch

REFINE_CAP = 0x100
BAD_INPUT = 0x200


def _run(collider1, collider2, mpr_tolerance, max_iterations, penetration):
    b = SceneBuilder()
    i1 = collider1.pack(b)
    i2 = collider2.pack(b)
    centers = np.zeros((2, 3))
    centers[i1] = collider1.center()
    centers[i2] = collider2.center()
    status, out = kernels.mpr_batch(b.build().to_device(), [i1], [i2], centers, mpr_tolerance,
                                    max_iterations, penetration)
    st = int(status.cpu().item())
    if st & BAD_INPUT:
        raise ValueError("invalid collider")
    if st & REFINE_CAP:
        raise RuntimeError("MPR portal refinement did not terminate within %d iterations"
                           % max_iterations)
    return st, (out.cpu().numpy()[0] if out is not None else None)


def mpr_intersection(collider1, collider2, mpr_tolerance=0.0001, max_iterations=100):
    """Intersection test with Minkowski Portal Refinement (mpr.py:20-49)."""
    st, _ = _run(collider1, collider2, mpr_tolerance, max_iterations, False)
    return bool(st & 1)


def mpr_penetration(collider1, collider2, mpr_tolerance=0.0001, max_iterations=100):
    """MPR with penetration info (mpr.py:52-109): ``(intersection, depth,
    penetration_direction, contact_position)``; the last three are None when
    the colliders do not intersect.  Adding ``depth * penetration_direction``
    to the second collider separates the two."""
    st, out = _run(collider1, collider2, mpr_tolerance, max_iterations, True)
    if not st & 1:
        return False, None, None, None
    return True, float(out[0]), out[1:4].copy(), out[4:7].copy()
