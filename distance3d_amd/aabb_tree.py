"""AABB tree with the reference's interface (aabb_tree.py:15-191, 465-527),
backed by a linear BVH built and queried on the GPU.

Leaf rows are numbered exactly as the reference numbers them (a bulk insert
into an empty tree puts leaves at rows 0..n-1, later inserts append n leaf
rows followed by n parent rows), so ``external_data_list`` and
``insert_index_list`` index the same way.  Overlap results are the
reference's sets; their order is ascending (the reference returns
depth-first order, which depends on its tree layout).
"""
import numpy as np

from . import kernels

INDEX_NONE = -1


class AabbTree:
    """AABB tree for broad phase collision detection."""

    def __init__(self):
        self.root = INDEX_NONE
        self.filled_len = 0
        self.aabbs = np.empty((0, 3, 2))
        self.external_data_list = []
        self.insert_index_list = []
        self.insert_index_max = 0
        self._leaf_rows = np.empty(0, dtype=np.int64)
        self._leaf_aabbs = np.empty((0, 3, 2))
        self._bvh = None

    def insert_aabbs(self, aabbs, external_data_list=None, pre_insertion_methode="none"):
        """Insert (n, 3, 2) aabbs; ``pre_insertion_methode`` only changes
        the reference's tree shape, not which leaves overlap, and is
        accepted for compatibility."""
        aabbs = np.asarray(aabbs, dtype=np.float64).reshape(-1, 3, 2)
        k = len(aabbs)
        if k == 0:
            return
        assert external_data_list is None or len(external_data_list) == k
        if pre_insertion_methode not in ("none", "sort", "shuffle"):
            raise ValueError("unknown pre_insertion_methode %r" % pre_insertion_methode)
        first = self.filled_len
        rows = np.arange(first, first + k)
        n_parents = k - 1 if first == 0 else k
        new_len = first + k + n_parents
        data = list(external_data_list) if external_data_list is not None else [None] * k
        self.external_data_list = (self.external_data_list[:first] + data
                                   + [None] * n_parents)
        self.insert_index_list = (self.insert_index_list[:first]
                                  + list(range(self.insert_index_max, self.insert_index_max + k))
                                  + [None] * n_parents)
        self.insert_index_max += k
        grown = np.zeros((new_len, 3, 2))
        grown[:first] = self.aabbs[:first]
        grown[first:first + k] = aabbs
        self.aabbs = grown
        self._leaf_rows = np.concatenate([self._leaf_rows, rows])
        self._leaf_aabbs = np.concatenate([self._leaf_aabbs, aabbs])
        self.filled_len = new_len
        self.root = 0 if first == 0 and k == 1 else (self.filled_len - 1)
        self._bvh = None

    def insert_aabb(self, aabb, external_data=None):
        self.insert_aabbs([aabb], [external_data], pre_insertion_methode="none")

    def _tree(self):
        if self._bvh is None:
            self._bvh = kernels.bvh_build(self._leaf_aabbs)
        return self._bvh

    def update_leaf_aabbs(self, aabbs):
        """New boxes for the leaves, in insertion order (same count): the
        device tree, when built, is refitted instead of rebuilt."""
        aabbs = np.asarray(aabbs, dtype=np.float64).reshape(-1, 3, 2)
        if len(aabbs) != len(self._leaf_rows):
            raise ValueError("expected %d boxes, got %d" % (len(self._leaf_rows), len(aabbs)))
        self._leaf_aabbs = aabbs.copy()
        self.aabbs[self._leaf_rows] = aabbs
        if self._bvh is not None:
            kernels.bvh_refit(self._bvh, self._leaf_aabbs)

    def overlaps_aabb(self, aabb):
        """(is_overlapping, leaf rows overlapping ``aabb``)."""
        if len(self._leaf_rows) == 0:
            return False, np.empty(0, dtype=np.int64)
        pairs = kernels.bvh_query(self._tree(), np.asarray(aabb, dtype=np.float64)[None])
        leaves = self._leaf_rows[pairs[:, 1].cpu().numpy()]
        return len(leaves) > 0, leaves

    def overlaps_aabb_tree(self, other):
        """(is_overlapping, unique self rows, unique other rows, pairs)."""
        if len(self._leaf_rows) == 0 or len(other._leaf_rows) == 0:
            e = np.empty(0, dtype=np.int64)
            return False, e, e, []
        pairs = kernels.bvh_query(self._tree(), other._leaf_aabbs).cpu().numpy()
        s = self._leaf_rows[pairs[:, 1]]
        o = other._leaf_rows[pairs[:, 0]]
        order = np.lexsort((o, s))
        s, o = s[order], o[order]
        return len(s) > 0, np.unique(s), np.unique(o), list(zip(s.tolist(), o.tolist()))

    def get_root_aabb(self):
        """Union of all leaf boxes (the reference's root box)."""
        b = self._leaf_aabbs
        return np.stack([b[:, :, 0].min(axis=0), b[:, :, 1].max(axis=0)], axis=1)


def all_aabbs_overlap(aabbs1, aabbs2):
    """Every (i, j) with aabb_overlap(aabbs1[i], aabbs2[j]), in row-major
    order (aabb_tree.py:465-500): (unique i, unique j, list of pairs)."""
    a1 = np.asarray(aabbs1, dtype=np.float64).reshape(-1, 3, 2)
    a2 = np.asarray(aabbs2, dtype=np.float64).reshape(-1, 3, 2)
    if len(a1) == 0 or len(a2) == 0:
        e = np.empty(0, dtype=np.int64)
        return e, e, []
    if len(a1) * len(a2) <= kernels.DENSE_OVERLAP_MAX_TESTS:
        pairs = kernels.overlap_dense(a1, a2).cpu().numpy().astype(np.int64)
    else:
        pairs = kernels.bvh_query(kernels.bvh_build(a2), a1).cpu().numpy().astype(np.int64)
    return np.unique(pairs[:, 0]), np.unique(pairs[:, 1]), list(map(tuple, pairs.tolist()))


def aabb_overlap(aabb1, aabb2):
    """Closed-interval overlap of two (3, 2) boxes (aabb_tree.py:503-527)."""
    return len(all_aabbs_overlap(np.asarray(aabb2)[None], np.asarray(aabb1)[None])[2]) > 0
