"""Self-collision detection of a BVH of colliders (self_collision.py:5-64):
broad phase on the GPU AABB tree, then one batched GJK intersection launch
over all candidate pairs.

The reference asks the tree once per collider (``aabb_overlapping_colliders``
with that collider's AABB, broad_phase.py:174-200) and runs one GJK
intersection test per candidate.  Here every collider is packed into one
scene, its AABBs come from one launch, the tree answers all of them in one
query and every candidate pair goes through one GJK launch: a whole
broad-to-narrow pass is three launches whatever the collider count.  The
reference's loop is then replayed on the host over those results, so the
outputs are its outputs: a frame is a candidate of itself unless its
whitelist holds it, a frame already marked is skipped, the first hit ends a
frame's candidates, the contacts dict is filled in the same order, and a
frame without a whitelist raises KeyError where the reference would.  A
frame's candidates come in ascending leaf order, the order
AabbTree.overlaps_aabb returns here (the reference's depth-first order
depends on its tree layout); with symmetric whitelists (each frame
whitelists another exactly when that one whitelists it back), which of
several hits ends a frame changes no contact flag.
"""
import numpy as np

from . import kernels
from .scene import SceneBuilder

_MISSING = object()


def _candidates(bvh):
    """(frames, per-frame candidate lists of frame indices, per-frame
    whitelists (_MISSING where the frame has none), GJK hit flags of the
    candidates in the same order).  A frame's candidates are the frames whose
    leaf boxes its AABB overlaps, in leaf order, each once (a dict in the
    reference), without those in its whitelist."""
    frames = list(bvh.colliders_.keys())
    whitelists = [bvh.self_collision_whitelists_.get(f, _MISSING) for f in frames]
    tree = bvh.aabbtree_
    if not frames or len(tree._leaf_rows) == 0:
        return frames, [[] for _ in frames], whitelists, np.zeros(0, dtype=bool)
    index = {f: i for i, f in enumerate(frames)}
    b = SceneBuilder()
    ids = [bvh.colliders_[f].pack(b) for f in frames]
    scene = b.build().to_device()
    aabbs = kernels.aabb_batch(scene, ids)
    # (query frame, leaf) pairs sorted by query, then leaf: per frame the leaf
    # order of AabbTree.overlaps_aabb
    hits = kernels.bvh_query(tree._tree(), aabbs).cpu().numpy()
    bounds = np.searchsorted(hits[:, 0], np.arange(len(frames) + 1))
    data = tree.external_data_list
    cand = []
    for q in range(len(frames)):
        found = dict.fromkeys(data[tree._leaf_rows[k]][0]
                              for k in hits[bounds[q]:bounds[q + 1], 1])
        wl = whitelists[q]
        for frame in (() if wl is _MISSING else wl):
            found.pop(frame, None)
        cand.append([index[f2] for f2 in found])
    pa = [ids[q] for q, c in enumerate(cand) for _ in c]
    pb = [ids[j] for c in cand for j in c]
    if not pa:
        return frames, cand, whitelists, np.zeros(0, dtype=bool)
    hit, _ = kernels.gjk_intersection_batch(scene, pa, pb)
    return frames, cand, whitelists, hit.cpu().numpy().astype(bool)


def detect(bvh):
    """Maps each collider frame to whether it touches another collider
    (self_collision.py:5-36)."""
    frames, cand, whitelists, hits = _candidates(bvh)
    contacts = {}
    pos = 0
    for q, f in enumerate(frames):
        c = cand[q]
        h = hits[pos:pos + len(c)]
        pos += len(c)
        if f in contacts:
            continue  # marked by an earlier frame's contact
        if whitelists[q] is _MISSING:
            raise KeyError(f)
        contacts[f] = False
        for j, hit in zip(c, h):
            if hit:
                contacts[f] = True
                contacts[frames[j]] = True
                break
    return contacts


def detect_any(bvh):
    """Is there any self collision? (self_collision.py:39-64)"""
    frames, cand, whitelists, hits = _candidates(bvh)
    pos = 0
    for q, f in enumerate(frames):
        if whitelists[q] is _MISSING:
            raise KeyError(f)
        n = len(cand[q])
        if np.any(hits[pos:pos + n]):
            return True
        pos += n
    return False
