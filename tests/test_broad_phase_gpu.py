"""GPU broad phase (LBVH build + stackless query) against the oracle's
brute-force closed-interval pair sets: the sets must be identical (bit-exact
FP64 predicate), for random, touching, duplicated and degenerate inputs, and
through the AabbTree / BoundingVolumeHierarchy drop-ins."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal

from distance3d_amd import aabb_tree, broad_phase, colliders, kernels, self_collision, workloads
from distance3d_amd.scene import SceneBuilder
from oracle import oracle
from tests import geometry_brute as gb

pytestmark = pytest.mark.gpu


def _boxes(rs, n, lo=0.0, hi=10.0, wmin=0.01, wmax=1.0):
    a = rs.uniform(lo, hi, (n, 3))
    return np.stack([a, a + rs.uniform(wmin, wmax, (n, 3))], axis=2)


def _gpu_pairs(boxes, queries):
    bvh = kernels.bvh_build(boxes)
    return kernels.bvh_query(bvh, queries).cpu().numpy().astype(np.int64)


def _oracle_pairs(boxes, queries):
    p, cnt = oracle.all_aabbs_overlap(queries, boxes)
    assert len(p) == cnt
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def test_random_sets_exact(gpu):
    rs = np.random.RandomState(1)
    boxes = _boxes(rs, 5000)
    queries = _boxes(rs, 1500, wmax=2.0)
    queries[:100, 0, 0] = boxes[:100, 0, 1]  # touching faces count (closed intervals)
    queries[100:200] = boxes[100:200]        # identical boxes
    g = _gpu_pairs(boxes, queries)
    o = _oracle_pairs(boxes, queries)
    assert np.array_equal(g, o)
    assert len(o) > 1000


def test_single_box_and_empty_query(gpu):
    box = np.array([[[0.0, 1.0], [0.0, 1.0], [0.0, 1.0]]])
    q = np.array([[[1.0, 2.0], [1.0, 2.0], [1.0, 2.0]], [[1.5, 2.0], [0.0, 1.0], [0.0, 1.0]]])
    assert _gpu_pairs(box, q).tolist() == [[0, 0]]
    bvh = kernels.bvh_build(box)
    assert kernels.bvh_query(bvh, np.zeros((0, 3, 2))).shape[0] == 0


def test_equal_morton_codes(gpu):
    """Many boxes share one centre: the hierarchy must tie-break by position."""
    rs = np.random.RandomState(2)
    n = 3000
    w = rs.uniform(0.001, 1.0, (n, 3))
    boxes = np.stack([5.0 - w, 5.0 + w], axis=2)
    boxes[:10] = _boxes(rs, 10)
    queries = _boxes(rs, 200, wmax=3.0)
    assert np.array_equal(_gpu_pairs(boxes, queries), _oracle_pairs(boxes, queries))


def test_broad_phase_scene_subset_exact(gpu):
    w = workloads.broad_phase_scene(seed=3, n_boxes=200000, n_capsules=20000)
    dev = w.scene.to_device()
    nb = w.meta["n_boxes"]
    ids = np.arange(w.scene.n_shapes, dtype=np.int32)
    aabbs = kernels.aabb_batch(dev, ids)
    # AABBs bit-identical to the checker's box_aabb / capsule_aabb
    assert np.array_equal(aabbs.cpu().numpy(), oracle.aabbs(w.scene, ids))
    bvh = kernels.bvh_build(aabbs[:nb])
    g = kernels.bvh_query(bvh, aabbs[nb:]).cpu().numpy().astype(np.int64)
    host = aabbs.cpu().numpy()
    o, cnt = oracle.overlap_pairs_sorted(host[nb:], host[:nb], cap=10 ** 6)
    o = o[np.lexsort((o[:, 1], o[:, 0]))]
    assert cnt == len(g)
    assert np.array_equal(g, o)
    # narrow phase on the candidates: capsule (query) vs box (leaf)
    r = kernels.gjk_distance_batch(dev, nb + g[:, 0], g[:, 1])
    od = oracle.gjk_distance(w.scene, nb + g[:, 0], g[:, 1])
    assert np.array_equal(r.dist.cpu().numpy() == 0.0, od["dist"] == 0.0)


# ------------------------------------------------ AabbTree drop-in

def _test_aabbs():  # test_aabb_tree.py:6-14
    aabbs = np.array([[[0, 2], [0, 2], [0, 1]], [[1, 3], [1, 3], [0, 1]],
                      [[3, 4], [0, 1], [0, 1]]], dtype=float)
    aabbs2 = np.array([[[1.5, 3.5], [0, 0.5], [0, 1]], [[1.5, 2], [2.5, 4], [0, 1]],
                       [[3.5, 4.5], [0, 0.5], [0, 1]]], dtype=float)
    return aabbs, aabbs2


def test_aabb_tree_creation(gpu):  # test_aabb_tree.py:17-36
    aabbs, aabbs2 = _test_aabbs()
    tree = aabb_tree.AabbTree()
    tree.insert_aabbs([])
    tree.insert_aabbs(aabbs, ["test"] * len(aabbs), pre_insertion_methode="shuffle")
    _, ov = tree.overlaps_aabb(aabbs2[0])
    ov = np.sort(ov)
    assert ov[0] == 0 and ov[1] == 2
    _, ov = tree.overlaps_aabb(aabbs2[1])
    assert np.sort(ov)[0] == 1
    _, ov = tree.overlaps_aabb(aabbs2[2])
    assert np.sort(ov)[0] == 2


def test_aabb_tree_overlap_with_other_tree(gpu):  # test_aabb_tree.py:39-55
    aabbs, aabbs2 = _test_aabbs()
    t1 = aabb_tree.AabbTree()
    t1.insert_aabbs(aabbs, ["1"] * 3)
    t2 = aabb_tree.AabbTree()
    t2.insert_aabbs(aabbs2, ["2"] * 3)
    hit, o1, o2, pairs = t1.overlaps_aabb_tree(t2)
    assert hit
    assert (o1 == np.array([0, 1, 2])).all()
    assert (o2 == np.array([0, 1, 2])).all()
    assert set(pairs) == {(0, 0), (2, 0), (1, 1), (2, 2)}


def test_aabb_tree_insert_single(gpu):  # test_aabb_tree.py:87-122
    a1 = np.array([[0, 10], [0, 10], [0, 10]], dtype=np.float64)
    a2 = np.array([[8, 12], [3, 4], [4, 5]], dtype=np.float64)
    a3 = np.array([[20, 30], [0, 1], [0, 2]], dtype=np.float64)
    a4 = np.array([[-50, 50], [-50, 50], [-50, 50]], dtype=np.float64)
    tree = aabb_tree.AabbTree()
    tree.insert_aabb(a1, 1)
    tree.insert_aabb(a2, 2)
    tree.insert_aabb(a3, 3)
    hit, ov = tree.overlaps_aabb(a1)
    ext = [tree.external_data_list[k] for k in ov]
    ins = [tree.insert_index_list[k] for k in ov]
    assert hit and 1 in ext and 2 in ext and 0 in ins and 1 in ins
    hit, ov = tree.overlaps_aabb(a4)
    ext = [tree.external_data_list[k] for k in ov]
    ins = [tree.insert_index_list[k] for k in ov]
    assert hit and sorted(ext) == [1, 2, 3] and sorted(ins) == [0, 1, 2]
    # rows follow the reference: leaf 0, then (leaf, parent) per insert
    assert tree.filled_len == 5
    assert list(ov) == [0, 1, 3]
    assert_array_almost_equal(tree.get_root_aabb(), [[0, 30], [0, 10], [0, 10]])


def test_all_aabbs_overlap_row_major(gpu):  # aabb_tree.py:465-500
    rs = np.random.RandomState(4)
    a, b = _boxes(rs, 300), _boxes(rs, 200)
    u1, u2, pairs = aabb_tree.all_aabbs_overlap(a, b)
    o, _ = oracle.all_aabbs_overlap(a, b)
    assert pairs == list(map(tuple, o.tolist()))
    assert np.array_equal(u1, np.unique(o[:, 0])) and np.array_equal(u2, np.unique(o[:, 1]))
    assert aabb_tree.aabb_overlap(a[o[0, 0]], b[o[0, 1]])


def test_overlap_dense_row_major(gpu):  # aabb_tree.py:465-500
    """Brute-force kernel: the oracle's row-major list exactly, for random,
    touching, identical, NaN and all-overlapping sets, sizes not multiples
    of the wavefront or row block, a cap below the count, and empty sets."""
    rs = np.random.RandomState(14)
    a, b = _boxes(rs, 1001), _boxes(rs, 777, wmax=2.0)
    a[:50, 0, 0] = b[:50, 0, 1]
    a[50:100] = b[50:100]
    a[100, 1, 0] = np.nan
    b[7, 2, 1] = np.nan
    o, _ = oracle.all_aabbs_overlap(a, b)
    assert np.array_equal(kernels.overlap_dense(a, b).cpu().numpy(), o)
    assert np.array_equal(kernels.overlap_dense(a, b, cap=10).cpu().numpy(), o)
    # everything overlaps everything (a shared point)
    c = np.zeros((701, 3, 2)); c[:, :, 0] = -rs.uniform(0.1, 1, (701, 3)); c[:, :, 1] = rs.uniform(0.1, 1, (701, 3))
    d = np.zeros((333, 3, 2)); d[:, :, 0] = -rs.uniform(0.1, 1, (333, 3)); d[:, :, 1] = rs.uniform(0.1, 1, (333, 3))
    g = kernels.overlap_dense(c, d).cpu().numpy()
    assert len(g) == 701 * 333
    assert np.array_equal(g, np.stack(np.meshgrid(np.arange(701), np.arange(333), indexing="ij"), -1).reshape(-1, 2))
    assert len(kernels.overlap_dense(a[:0], b)) == 0 and len(kernels.overlap_dense(a, b[:0])) == 0


# ---------------------------------- BoundingVolumeHierarchy drop-in

class _TM:
    """Minimal transform manager: poses of frames in "origin"."""

    def __init__(self):
        self.T = {}

    def add_transform(self, a, b, A2B):
        assert b == "origin"
        self.T[a] = np.asarray(A2B, dtype=float)

    def get_transform(self, a, b):
        return self.T[a]


def _robotish(rs, n=30):
    tm = _TM()
    bvh = broad_phase.BoundingVolumeHierarchy(tm, "base")
    for k in range(n):
        T = gb.random_pose(rs, 0.8)
        tm.add_transform("f%d" % k, "origin", T)
        c = [colliders.Sphere(T[:3, 3], rs.uniform(0.1, 0.4)),
             colliders.Box(T, rs.uniform(0.1, 0.6, 3)),
             colliders.Capsule(T, rs.uniform(0.05, 0.2), rs.uniform(0.2, 0.8))][k % 3]
        bvh.add_collider("f%d" % k, c)
    return tm, bvh


def test_bvh_with_self_matches_other_bvh(gpu):  # test_broad_phase.py:32-101
    rs = np.random.RandomState(5)
    tm, bvh = _robotish(rs)
    self_pairs = bvh.aabb_overlapping_with_self()
    other = bvh.aabb_overlapping_with_other_bvh(bvh)
    key = lambda p: (p[0][0], p[1][0])
    assert sorted(map(key, self_pairs)) == sorted(key(p) for p in other if p[0][0] != p[1][0])
    # pose update rebuilds the tree; same poses give the same pairs
    bvh.update_collider_poses()
    assert sorted(map(key, bvh.aabb_overlapping_with_self())) == sorted(map(key, self_pairs))
    # every pair from the tree overlaps by brute force, and vice versa
    frames = list(bvh.colliders_)
    boxes = np.array([bvh.colliders_[f].aabb() for f in frames])
    o, _ = oracle.all_aabbs_overlap(boxes, boxes)
    expect = sorted((frames[i], frames[j]) for i, j in o.tolist() if i != j)
    assert sorted(map(key, self_pairs)) == expect


def test_bvh_refit_matches_brute_force(gpu):
    """Refitting a built tree with moved, grown and shrunk boxes gives the
    brute-force pair sets; two refits in a row, and trees of 1 and 2 boxes."""
    rs = np.random.RandomState(11)
    for n in (1, 2, 3000):
        boxes = _boxes(rs, n)
        bvh = kernels.bvh_build(boxes)
        queries = _boxes(rs, 500, wmax=2.0)
        for step in range(2):
            moved = boxes + rs.uniform(-3.0, 3.0, (n, 3, 1))
            moved[:, :, 1] += rs.uniform(0.0, 1.0, (n, 3)) * (step == 0)
            kernels.bvh_refit(bvh, moved)
            got = kernels.bvh_query(bvh, queries).cpu().numpy().astype(np.int64)
            assert np.array_equal(got, _oracle_pairs(moved, queries))
            assert np.array_equal(got, _gpu_pairs(moved, queries))


def test_update_collider_poses_refits(gpu):
    """Moving frames and updating poses keeps the tree object and refits it;
    overlaps equal brute force over the new boxes."""
    rs = np.random.RandomState(12)
    tm, bvh = _robotish(rs, 40)
    bvh.aabb_overlapping_with_self()  # builds the device tree
    tree, dev = bvh.aabbtree_, bvh.aabbtree_._bvh
    assert dev is not None
    for f in list(bvh.colliders_):
        T = tm.get_transform(f, "origin").copy()
        T[:3, 3] += rs.uniform(-0.5, 0.5, 3)
        tm.add_transform(f, "origin", T)
    bvh.update_collider_poses()
    assert bvh.aabbtree_ is tree and bvh.aabbtree_._bvh is dev
    frames = list(bvh.colliders_)
    boxes = np.array([bvh.colliders_[f].aabb() for f in frames])
    o, _ = oracle.all_aabbs_overlap(boxes, boxes)
    expect = sorted((frames[i], frames[j]) for i, j in o.tolist() if i != j)
    key = lambda p: (p[0][0], p[1][0])
    assert sorted(map(key, bvh.aabb_overlapping_with_self())) == expect


def test_self_collision_detect(gpu):  # self_collision.py:5-64
    """The batched pass against the reference's per-frame loop run here with
    the drop-in's per-frame tree query (aabb_overlapping_colliders) and the
    oracle's GJK intersection: whitelists holding the frame and its chain
    neighbours, some one-sided, one frame left its own candidate."""
    rs = np.random.RandomState(6)
    tm, bvh = _robotish(rs, 24)
    frames = list(bvh.colliders_)
    b = SceneBuilder()
    ids = {f: bvh.colliders_[f].pack(b) for f in frames}
    host = b.build()

    def intersects(f, f2):
        hit, _ = oracle.gjk_intersection(host, [ids[f]], [ids[f2]])
        return bool(hit[0])

    def reference_detect():
        contacts = {}
        for f, c in bvh.colliders_.items():
            if f in contacts:
                continue
            cand = bvh.aabb_overlapping_colliders(c, whitelist=bvh.self_collision_whitelists_[f])
            contacts[f] = False
            for f2 in cand:
                if intersects(f, f2):
                    contacts[f] = contacts[f2] = True
                    break
        return contacts

    def reference_detect_any():
        for f, c in bvh.colliders_.items():
            cand = bvh.aabb_overlapping_colliders(c, whitelist=bvh.self_collision_whitelists_[f])
            if any(intersects(f, f2) for f2 in cand):
                return True
        return False

    for k, f in enumerate(frames):
        wl = [f]
        if k + 1 < len(frames):
            wl.append(frames[k + 1])
        if k > 0 and k % 3 != 0:  # every third frame does not whitelist its predecessor
            wl.append(frames[k - 1])
        bvh.self_collision_whitelists_[f] = tuple(wl)
    for variant in range(2):
        want = reference_detect()
        got = self_collision.detect(bvh)
        assert got == want and list(got) == list(want), variant
        assert 0 < sum(want.values()) < len(frames)
        assert self_collision.detect_any(bvh) == reference_detect_any()
        if variant == 0:
            # a frame that is its own candidate touches itself
            quiet = next(f for f in frames if not got[f])
            bvh.self_collision_whitelists_[quiet] = ()
    assert got[quiet]
    del bvh.self_collision_whitelists_[frames[0]]
    with pytest.raises(KeyError):
        self_collision.detect(bvh)


def test_cube_sphere_unique_pairs_edge_case(gpu):  # test_cube_unique_pairs_edge_case.py:8-31
    """Exactly 162 tetrahedron-box pairs, through both the brute-force
    overlap kernel (all_aabbs_overlap) and the tree-vs-tree query
    (AabbTree.overlaps_aabb_tree), equal to the oracle's set."""
    from tests.test_oracle_pins import cube_sphere_tetrahedron_boxes
    rb1, rb2 = cube_sphere_tetrahedron_boxes()
    # tetrahedron boxes computed on the GPU, bit-equal to the oracle's
    a1, a2 = rb1.aabbs, rb2.aabbs
    assert np.array_equal(a1, oracle.tet_mesh_prep(rb1.tetrahedra_points)[0])
    assert np.array_equal(a2, oracle.tet_mesh_prep(rb2.tetrahedra_points)[0])
    want, cnt = oracle.all_aabbs_overlap(a1, a2)
    assert cnt == 162
    _, _, dense = aabb_tree.all_aabbs_overlap(a1, a2)
    assert len(dense) == 162
    assert [tuple(p) for p in want.tolist()] == dense  # row-major order
    _, u1, u2, tree = rb1.aabb_tree.overlaps_aabb_tree(rb2.aabb_tree)
    assert len(tree) == 162
    assert set(tree) == set(dense)
    assert np.array_equal(u1, np.unique(want[:, 0]))
    assert np.array_equal(u2, np.unique(want[:, 1]))
