"""Self-collision detection of a BVH of colliders (self_collision.py:5-64):
broad phase on the GPU AABB tree, then one batched GJK intersection launch
over all candidate pairs."""
import numpy as np

from . import kernels
from .scene import SceneBuilder


def _candidates(bvh):
    frames = list(bvh.colliders_.keys())
    index = {f: i for i, f in enumerate(frames)}
    pairs = []
    for f in frames:
        wl = bvh.self_collision_whitelists_.get(f, ())
        for f2 in bvh.aabb_overlapping_colliders(bvh.colliders_[f], whitelist=wl):
            if f2 != f:
                pairs.append((index[f], index[f2]))
    return frames, pairs


def _hits(bvh, frames, pairs):
    if not pairs:
        return np.zeros(0, dtype=bool)
    b = SceneBuilder()
    ids = [bvh.colliders_[f].pack(b) for f in frames]
    scene = b.build().to_device()
    pa = [ids[p[0]] for p in pairs]
    pb = [ids[p[1]] for p in pairs]
    hit, _ = kernels.gjk_intersection_batch(scene, pa, pb)
    return hit.cpu().numpy().astype(bool)


def detect(bvh):
    """Maps each collider frame to whether it touches another collider."""
    frames, pairs = _candidates(bvh)
    hits = _hits(bvh, frames, pairs)
    contacts = {f: False for f in frames}
    for (i, j), h in zip(pairs, hits):
        if h:
            contacts[frames[i]] = True
            contacts[frames[j]] = True
    return contacts


def detect_any(bvh):
    """Is there any self collision?"""
    frames, pairs = _candidates(bvh)
    return bool(np.any(_hits(bvh, frames, pairs)))
