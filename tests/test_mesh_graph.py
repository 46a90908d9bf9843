"""MeshGraph support: the reference climbs the mesh graph to an extreme
vertex (mesh.py:12-139, colliders.py:187-240); the library evaluates the
argmax over the mesh's vertices (a posed hull).  test_mesh.py:7-19 pins the
two as equal on random convex meshes; these tests restate that check against
the oracle's argmax (CPU) and the GPU support (GPU)."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal
from scipy.spatial import ConvexHull

from distance3d_amd import colliders
from distance3d_amd.scene import SceneBuilder
from oracle import oracle

# mesh.py:9 (10 * utils.EPSILON, utils.py:8)
PROJECTION_LENGTH_EPSILON = 10.0 * np.finfo(float).eps


def random_convex_mesh(rs, n_vertices=10):
    """Points on an ellipsoid surface, their convex hull's triangles and a
    random pose (the shape of random.randn_convex, random.py:407-460)."""
    while True:
        d = rs.randn(n_vertices, 3)
        d /= np.linalg.norm(d, axis=1)[:, np.newaxis]
        vertices = d * (1.0 + rs.rand(3))[np.newaxis]
        hull = ConvexHull(vertices)
        if len(hull.vertices) == n_vertices:
            break
    q = rs.randn(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    mesh2origin = np.eye(4)
    mesh2origin[:3, :3] = R
    mesh2origin[:3, 3] = rs.randn(3)
    return mesh2origin, vertices, hull.simplices


class HillClimb:
    """Restatement of MeshHillClimbingSupportFunction (mesh.py:12-88) and
    hill_climb_mesh_extreme (mesh.py:90-139): six axis-extreme shortcuts,
    then climbing to any neighbour that projects more than
    PROJECTION_LENGTH_EPSILON further, starting from the previous result."""

    def __init__(self, mesh2origin, vertices, triangles):
        self.mesh2origin = mesh2origin
        self.vertices = vertices
        self.first_idx = int(np.min(triangles))
        self.connections = {}
        for i, j, k in triangles:
            for a, b, c in ((i, j, k), (j, i, k), (k, i, j)):
                self.connections.setdefault(int(a), set()).update((int(b), int(c)))
        self.shortcuts = [np.argmax(vertices[:, 0]), np.argmax(vertices[:, 1]),
                          np.argmax(vertices[:, 2]), np.argmin(vertices[:, 0]),
                          np.argmin(vertices[:, 1]), np.argmin(vertices[:, 2])]

    def __call__(self, search_direction):
        d = self.mesh2origin[:3, :3].T.dot(search_direction)
        best = self.first_idx
        for idx in self.shortcuts:
            if d.dot(self.vertices[idx] - self.vertices[best]) > PROJECTION_LENGTH_EPSILON:
                best = int(idx)
        converged = False
        while not converged:
            converged = True
            for idx in self.connections[best]:
                if d.dot(self.vertices[idx] - self.vertices[best]) > PROJECTION_LENGTH_EPSILON:
                    best = idx
                    converged = False
        self.first_idx = best
        return best, self.mesh2origin[:3, 3] + self.mesh2origin[:3, :3].dot(self.vertices[best])


def _meshes(seed, n_meshes=10):
    rs = np.random.RandomState(seed)
    out = []
    for _ in range(n_meshes):
        mesh2origin, vertices, triangles = random_convex_mesh(rs)
        dirs = rs.randn(20, 3)
        dirs /= np.linalg.norm(dirs, axis=1)[:, np.newaxis]
        out.append((mesh2origin, vertices, triangles, dirs))
    return out


def test_oracle_argmax_equals_hill_climb():
    """The oracle's posed-hull support of a MeshGraph reaches the vertex the
    reference's hill climb reaches (test_mesh.py:7-19's check)."""
    for mesh2origin, vertices, triangles, dirs in _meshes(2323):
        climb = HillClimb(mesh2origin, vertices, triangles)
        b = SceneBuilder()
        colliders.MeshGraph(mesh2origin, vertices, triangles).pack(b)
        host = b.build()
        got = oracle.support(host, np.zeros(len(dirs), dtype=np.int32), dirs)
        for d, p in zip(dirs, got):
            idx, point = climb(d)
            assert idx == int(np.argmax(vertices @ (mesh2origin[:3, :3].T @ d)))
            assert_array_almost_equal(p, point)


def test_first_vertex_is_not_the_support():
    """The comparison above discriminates: a support that stayed at the
    climb's starting vertex would disagree with the argmax."""
    mesh2origin, vertices, triangles, dirs = _meshes(7, 1)[0]
    b = SceneBuilder()
    colliders.MeshGraph(mesh2origin, vertices, triangles).pack(b)
    host = b.build()
    got = oracle.support(host, np.zeros(len(dirs), dtype=np.int32), dirs)
    first = mesh2origin[:3, 3] + mesh2origin[:3, :3].dot(vertices[0])
    assert any(not np.allclose(p, first) for p in got)


@pytest.mark.gpu
def test_mesh_graph_support_gpu(gpu):
    """MeshGraph.support_function on the GPU: bit-identical to the oracle and
    equal to the reference's hill-climbed vertex."""
    for mesh2origin, vertices, triangles, dirs in _meshes(4242):
        mg = colliders.MeshGraph(mesh2origin, vertices, triangles)
        climb = HillClimb(mesh2origin, vertices, triangles)
        b = SceneBuilder()
        mg.pack(b)
        host = b.build()
        expect = oracle.support(host, np.zeros(len(dirs), dtype=np.int32), dirs)
        for d, e in zip(dirs, expect):
            p = np.asarray(mg.support_function(d))
            assert np.array_equal(p, e)
            assert_array_almost_equal(p, climb(d)[1])
