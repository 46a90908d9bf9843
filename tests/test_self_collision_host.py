"""Self-collision drop-in on the CPU: the batched pass (one AABB launch, one
tree query, one GJK launch, then a host replay) gives the outputs of the
reference's per-frame loop (self_collision.py:5-64, broad_phase.py:174-200)
for random overlap and hit tables, whitelists that leave a frame its own
candidate, asymmetric whitelists, frames held by two leaves and frames
without a whitelist.  The three launches are replaced by table lookups."""
import numpy as np
import pytest
import torch

from distance3d_amd import self_collision


class _Collider:
    def __init__(self, k):
        self.k = k

    def pack(self, builder):
        builder.ids.append(self.k)
        return self.k


class _Builder:
    def __init__(self):
        self.ids = []

    def build(self):
        return self

    def to_device(self):
        return self


class _Tree:
    def __init__(self, leaf_frames):
        self.external_data_list = [(f, None) for f in leaf_frames]
        self._leaf_rows = np.arange(len(leaf_frames))

    def _tree(self):
        return self


class _BVH:
    def __init__(self, frames, leaf_frames, whitelists):
        self.colliders_ = {f: _Collider(i) for i, f in enumerate(frames)}
        self.aabbtree_ = _Tree(leaf_frames)
        self.self_collision_whitelists_ = whitelists


def reference_detect(bvh, overlapping, intersects):
    """The reference's detect loop over an overlap table (frame -> frames of
    the overlapping leaves, in leaf order) and a GJK outcome table."""
    contacts = {}
    for frame in bvh.colliders_:
        if frame in contacts:
            continue
        candidates = dict.fromkeys(overlapping[frame])
        for w in bvh.self_collision_whitelists_[frame]:
            candidates.pop(w, None)
        contacts[frame] = False
        for frame2 in candidates:
            if intersects(frame, frame2):
                contacts[frame] = True
                contacts[frame2] = True
                break
    return contacts


def reference_detect_any(bvh, overlapping, intersects):
    for frame in bvh.colliders_:
        candidates = dict.fromkeys(overlapping[frame])
        for w in bvh.self_collision_whitelists_[frame]:
            candidates.pop(w, None)
        for frame2 in candidates:
            if intersects(frame, frame2):
                return True
    return False


def _outcome(fn, *args):
    try:
        return ("ok", fn(*args))
    except KeyError as e:
        return ("KeyError", e.args)


def test_batched_pass_replays_the_reference_loop(monkeypatch):
    rs = np.random.RandomState(0)
    checked = 0
    for trial in range(400):
        n = int(rs.randint(1, 10))
        frames = ["f%d" % k for k in range(n)]
        # leaves: every frame once, a few twice
        leaf_frames = frames + [frames[k] for k in rs.randint(0, n, rs.randint(0, 3))]
        rs.shuffle(leaf_frames)
        # leaf boxes overlap symmetrically; a frame's box overlaps its own leaves
        lo = np.triu(rs.rand(n, n) < 0.4, 1)
        ov = lo | lo.T | np.eye(n, dtype=bool)
        hit = rs.rand(n, n) < 0.3
        hit = np.triu(hit, 1) | np.triu(hit, 1).T | np.eye(n, dtype=bool)  # self: always
        whitelists = {}
        for k, f in enumerate(frames):
            if rs.rand() < 0.05:
                continue  # no whitelist: KeyError where the reference reads it
            wl = [frames[j] for j in range(n) if rs.rand() < 0.2]
            if rs.rand() < 0.7:
                wl.append(f)
            whitelists[f] = tuple(wl)
        bvh = _BVH(frames, leaf_frames, whitelists)
        idx = {f: i for i, f in enumerate(frames)}
        overlapping = {f: [g for g in leaf_frames if ov[idx[f], idx[g]]] for f in frames}
        intersects = lambda a, b: bool(hit[idx[a], idx[b]])

        def bvh_query(tree, queries):
            rows = [(q, k) for q in queries for k, g in enumerate(leaf_frames)
                    if ov[q, idx[g]]]
            return torch.tensor(rows, dtype=torch.int32).reshape(-1, 2)

        def gjk_intersection_batch(scene, pa, pb):
            return torch.tensor([int(hit[a, b]) for a, b in zip(pa, pb)], dtype=torch.uint8), None

        monkeypatch.setattr(self_collision, "SceneBuilder", _Builder)
        monkeypatch.setattr(self_collision.kernels, "aabb_batch", lambda scene, ids: list(ids))
        monkeypatch.setattr(self_collision.kernels, "bvh_query", bvh_query)
        monkeypatch.setattr(self_collision.kernels, "gjk_intersection_batch",
                            gjk_intersection_batch)
        want = _outcome(reference_detect, bvh, overlapping, intersects)
        got = _outcome(self_collision.detect, bvh)
        assert got == want, (trial, got, want)
        if want[0] == "ok":
            assert list(got[1]) == list(want[1])  # same insertion order
            checked += 1
        assert (_outcome(self_collision.detect_any, bvh)
                == _outcome(reference_detect_any, bvh, overlapping, intersects)), trial
    assert checked > 200
