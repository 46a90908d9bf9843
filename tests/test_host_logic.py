"""Host-side logic that needs no GPU: scene packing, work classes, reference
row numbering of the AABB tree."""
import numpy as np
import pytest

from distance3d_amd import autotune, colliders, workloads
from distance3d_amd.scene import SceneBuilder, hull_bound, plane_basis_from_normal


def test_work_class_boundaries():
    n = np.array([0, 1, 32, 33, 128, 129, 512, 513, 1024, 1025, 2048, 2049, 4096, 4097, 8192])
    assert autotune.work_class(n).tolist() == [0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7]


def test_hull_packing_fp32_exact_and_fp64_pool():
    b = SceneBuilder()
    k0 = b.add_hull(np.array([[0.5, 0.25, -1.0], [2.0, 3.0, 4.0]]))
    k1 = b.add_hull(np.array([[0.1, 0.2, 0.3]]))  # not FP32-exact
    s = b.build()
    assert s.shape_i32[k0].tolist() == [0, 2, 0, -1]
    assert s.shape_i32[k1].tolist() == [0, 1, 2, 0]
    assert np.array_equal(s.v64[0], [0.1, 0.2, 0.3])
    assert s.shape_f64[k0, 12] >= 9.0 and s.shape_f64[k0, 12] == hull_bound([[2, 3, 4]])


def test_collider_packing_layout():
    T = np.eye(4)
    T[:3, 3] = [1, 2, 3]
    b = SceneBuilder()
    colliders.Capsule(T, 0.5, 2.0).pack(b)
    colliders.Margin(colliders.Box(T, np.array([1.0, 2.0, 3.0])), 0.25).pack(b)
    colliders.Disk(np.zeros(3), 0.5, np.array([0.0, 0.0, 1.0])).pack(b)
    s = b.build()
    assert s.shape_i32[:, 0].tolist() == [2, 5, 7]
    assert s.shape_f64[0, 9:15].tolist() == [1, 2, 3, 0.5, 2.0, 0.0]
    assert s.shape_f64[1, 15] == 0.25
    x, y = plane_basis_from_normal(np.array([0.0, 0.0, 1.0]))
    assert np.allclose(s.shape_f64[2, :9].reshape(3, 3), np.column_stack((x, y, [0, 0, 1])))


def test_invalid_scene_rejected():
    b = SceneBuilder()
    b.add_hull(np.zeros((3, 3)))
    s = b.build()
    s.shape_i32[0, 1] = 99  # vertex range beyond the pool
    with pytest.raises(ValueError):
        s.validate()


def test_workload_generators_shapes():
    w = workloads.mixed_pairs(n_pairs=1000)
    assert w.n_pairs == 1000 and w.scene.n_shapes == 2000
    assert set(np.unique(w.scene.shape_i32[:, 0])) <= {0, 1, 2, 3, 4, 5}
    h = w.scene.shape_i32[w.scene.shape_i32[:, 0] == 0]
    assert h[:, 1].min() >= 8 and h[:, 1].max() <= 4096
    e = workloads.epa_pairs(n_pairs=100)
    assert e.scene.shape_i32[:, 1].min() >= 64 and e.scene.shape_i32[:, 1].max() <= 256
    scene, poses, reqs = workloads.serving_workload(n_requests=50, n_colliders=30)
    assert poses.shape[1] == 12 and all(1 <= r.n_pairs <= 4096 for r in reqs)
    assert all(np.all(r.pair_a != r.pair_b) for r in reqs)
    inst = workloads.posed_instances(scene, poses, reqs[0].pair_a, reqs[0].pair_b,
                                     reqs[0].pose_a, reqs[0].pose_b)
    assert inst.n_shapes == 2 * reqs[0].n_pairs
    assert np.array_equal(inst.shape_f64[0, :12], poses[reqs[0].pose_a[0]])
