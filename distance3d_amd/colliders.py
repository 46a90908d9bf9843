"""Colliders with the reference's constructors and methods (colliders.py:17-659).

Each collider packs itself into a scene (:mod:`distance3d_amd.scene`);
``support_function`` and ``aabb`` evaluate on the GPU through libd3d_hip.
Attribute names and the state they hold follow the reference so code written
against ``distance3d.colliders`` keeps working.
"""
import abc

import numpy as np

from . import kernels
from .scene import SceneBuilder, plane_basis_from_normal


class ConvexCollider(abc.ABC):
    """Convex collider base class (colliders.py:17-106)."""

    def __init__(self, artist=None):
        self.artist_ = artist

    @abc.abstractmethod
    def make_artist(self, c=None):
        """Make artist that represents this collider."""

    @abc.abstractmethod
    def first_vertex(self):
        """Vertex used to initialise GJK variants that need one."""

    @abc.abstractmethod
    def center(self):
        """(Approximate) centre of the collider."""

    @abc.abstractmethod
    def update_pose(self, pose):
        """Update the pose of the collider."""

    @abc.abstractmethod
    def collider2origin(self):
        """Pose of the collider, shape (4, 4)."""

    @abc.abstractmethod
    def pack(self, builder, margin=0.0):
        """Add this collider to a SceneBuilder; returns its shape index."""

    def _single(self):
        b = SceneBuilder()
        self.pack(b)
        return b.build().to_device()

    def support_function(self, search_direction):
        """Extreme point of the collider along ``search_direction``."""
        d = np.ascontiguousarray(search_direction, dtype=np.float64).reshape(1, 3)
        out = kernels.support_batch(self._single(), np.zeros(1, dtype=np.int32), d)
        return out.cpu().numpy()[0]

    def aabb(self):
        """Axis-aligned bounding box, shape (3, 2) = [axis][min, max]."""
        out = kernels.aabb_batch(self._single(), np.zeros(1, dtype=np.int32))
        return out.cpu().numpy()[0]


def _pose(p):
    return np.asarray(p, dtype=np.float64)


class ConvexHullVertices(ConvexCollider):
    """Convex hull of a set of world-frame vertices (colliders.py:109-144)."""

    def __init__(self, vertices, artist=None):
        super(ConvexHullVertices, self).__init__(artist)
        self.vertices = vertices

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.PointCollection3D(self.vertices, s=0.005, c=c)

    def first_vertex(self):
        return self.vertices[0]

    def center(self):
        return np.mean(self.vertices, axis=0)

    def update_pose(self, pose):
        raise NotImplementedError("update_pose is not implemented!")

    def collider2origin(self):
        return np.eye(4)

    def pack(self, builder, margin=0.0):
        return builder.add_hull(self.vertices, margin=margin)


class Box(ConvexHullVertices):
    """Oriented box (colliders.py:147-184).  ``vertices`` holds the eight
    corners like the reference; the kernels use the closed form of the same
    argmax."""

    def __init__(self, box2origin, size, artist=None):
        super(Box, self).__init__(_box_vertices(box2origin, size), artist)
        self.box2origin = box2origin
        self.size = size

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Box(size=self.size, A2B=self.box2origin, c=c)

    def center(self):
        return self.box2origin[:3, 3]

    def update_pose(self, pose):
        self.box2origin = pose
        self.vertices = _box_vertices(pose, self.size)
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        return self.box2origin

    def pack(self, builder, margin=0.0):
        return builder.add_box(_pose(self.box2origin), self.size, margin=margin)


# BOX_COORDS of geometry.py:135: itertools.product([-0.5, 0.5], repeat=3)
_BOX_COORDS = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5)
                        for z in (-0.5, 0.5)])


def _box_vertices(box2origin, size):
    """The eight corners t + (BOX_COORDS * size) R^T (geometry.py:138-157);
    host-side attribute only, the kernels never read it."""
    box2origin = _pose(box2origin)
    return box2origin[:3, 3] + (_BOX_COORDS * np.asarray(size)).dot(box2origin[:3, :3].T)


class MeshGraph(ConvexCollider):
    """Convex mesh (colliders.py:187-240).  The reference climbs the mesh
    graph to the extreme vertex; on a convex mesh that is the argmax over its
    vertices, which is what the kernels compute."""

    def __init__(self, mesh2origin, vertices, triangles, artist=None):
        super(MeshGraph, self).__init__(artist)
        self.mesh2origin = mesh2origin
        self.vertices = vertices
        self.triangles = triangles

    def make_artist(self, c=None):
        raise ImportError("MeshGraph artists need distance3d's visualization module")

    def first_vertex(self):
        return self.mesh2origin[:3, 3] + np.dot(self.mesh2origin[:3, :3], self.vertices[0])

    def center(self):
        return self.mesh2origin[:3, 3] + np.dot(
            self.mesh2origin[:3, :3], np.mean(self.vertices, axis=0))

    def update_pose(self, mesh2origin):
        self.mesh2origin = mesh2origin
        if self.artist_ is not None:
            self.artist_.set_data(mesh2origin)

    def collider2origin(self):
        return self.mesh2origin

    def pack(self, builder, margin=0.0):
        pose = _pose(self.mesh2origin)
        return builder.add_hull(self.vertices, pose[:3, :3], pose[:3, 3], margin=margin)


class Sphere(ConvexCollider):
    """Sphere (colliders.py:243-287)."""

    def __init__(self, center, radius, artist=None):
        super(Sphere, self).__init__(artist)
        self.c = center
        self.radius = radius

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Sphere(radius=self.radius, A2B=self.collider2origin(), c=c)

    def center(self):
        return self.c

    def first_vertex(self):
        return self.c + np.array([0, 0, self.radius], dtype=float)

    def update_pose(self, pose):
        self.c = pose[:3, 3]
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        sphere2origin = np.eye(4)
        sphere2origin[:3, 3] = self.c
        return sphere2origin

    def pack(self, builder, margin=0.0):
        return builder.add_sphere(self.c, self.radius, margin=margin)


class Capsule(ConvexCollider):
    """Capsule (colliders.py:290-340)."""

    def __init__(self, capsule2origin, radius, height, artist=None):
        super(Capsule, self).__init__(artist)
        self.capsule2origin = capsule2origin
        self.radius = radius
        self.height = height

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Capsule(height=self.height, radius=self.radius,
                                  A2B=self.capsule2origin, c=c)

    def center(self):
        return self.capsule2origin[:3, 3]

    def first_vertex(self):
        return self.capsule2origin[:3, 3] - (
            self.radius + 0.5 * self.height) * self.capsule2origin[:3, 2]

    def update_pose(self, pose):
        self.capsule2origin = pose
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        return self.capsule2origin

    def pack(self, builder, margin=0.0):
        return builder.add_capsule(self.capsule2origin, self.radius, self.height, margin=margin)


class Ellipsoid(ConvexCollider):
    """Ellipsoid (colliders.py:343-387)."""

    def __init__(self, ellipsoid2origin, radii, artist=None):
        super(Ellipsoid, self).__init__(artist)
        self.ellipsoid2origin = ellipsoid2origin
        self.radii = radii

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Ellipsoid(radii=self.radii, A2B=self.ellipsoid2origin, c=c)

    def center(self):
        return self.ellipsoid2origin[:3, 3]

    def first_vertex(self):
        return self.ellipsoid2origin[:3, 3] + self.radii[2] * self.ellipsoid2origin[:3, 2]

    def update_pose(self, pose):
        self.ellipsoid2origin = pose
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        return self.ellipsoid2origin

    def pack(self, builder, margin=0.0):
        return builder.add_ellipsoid(self.ellipsoid2origin, self.radii, margin=margin)


class Cylinder(ConvexCollider):
    """Cylinder (colliders.py:390-440)."""

    def __init__(self, cylinder2origin, radius, length, artist=None):
        super(Cylinder, self).__init__(artist)
        self.cylinder2origin = cylinder2origin
        self.radius = radius
        self.length = length

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Cylinder(length=self.length, radius=self.radius,
                                   A2B=self.cylinder2origin, c=c)

    def center(self):
        return self.cylinder2origin[:3, 3]

    def first_vertex(self):
        return self.cylinder2origin[:3, 3] + 0.5 * self.length * self.cylinder2origin[:3, 2]

    def update_pose(self, pose):
        self.cylinder2origin = pose
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        return self.cylinder2origin

    def pack(self, builder, margin=0.0):
        return builder.add_cylinder(self.cylinder2origin, self.radius, self.length, margin=margin)


class Disk(ConvexCollider):
    """Disk (colliders.py:443-497)."""

    def __init__(self, center, radius, normal, artist=None):
        super(Disk, self).__init__(artist)
        self.c = center
        self.radius = radius
        self.normal = normal

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Cylinder(A2B=self.collider2origin(), radius=self.radius,
                                   length=0.01 * self.radius, c=c)

    def center(self):
        return self.c

    def first_vertex(self):
        x, _ = plane_basis_from_normal(self.normal)
        return self.c + self.radius * x

    def update_pose(self, pose):
        self.c = pose[:3, 3]
        self.normal = pose[:3, 2]
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        x, y = plane_basis_from_normal(self.normal)
        disk2origin = np.eye(4)
        disk2origin[:3, :3] = np.column_stack((x, y, self.normal))
        disk2origin[:3, 3] = self.c
        return disk2origin

    def pack(self, builder, margin=0.0):
        return builder.add_disk(self.c, self.radius, self.normal, margin=margin)


class Ellipse(ConvexCollider):
    """Ellipse (colliders.py:500-551)."""

    def __init__(self, center, axes, radii, artist=None):
        super(Ellipse, self).__init__(artist)
        self.c = center
        self.axes = axes
        self.radii = radii

    def make_artist(self, c=None):
        raise ImportError("Ellipse artists need distance3d's visualization module")

    def center(self):
        return self.c

    def first_vertex(self):
        return self.c + self.axes[0] * self.radii[0]

    def update_pose(self, pose):
        self.c = pose[:3, 3]
        self.axes = pose[:3, :2].T
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        ellipse2origin = np.eye(4)
        ellipse2origin[:3, :2] = np.asarray(self.axes).T
        ellipse2origin[:3, 2] = np.cross(self.axes[0], self.axes[1])
        ellipse2origin[:3, 3] = self.c
        return ellipse2origin

    def pack(self, builder, margin=0.0):
        return builder.add_ellipse(self.c, self.axes, self.radii, margin=margin)


class Cone(ConvexCollider):
    """Cone (colliders.py:554-603)."""

    def __init__(self, cone2origin, radius, height, artist=None):
        super(Cone, self).__init__(artist)
        self.cone2origin = cone2origin
        self.radius = radius
        self.height = height

    def make_artist(self, c=None):
        import pytransform3d.visualizer as pv
        self.artist_ = pv.Cone(A2B=self.cone2origin, radius=self.radius,
                               height=self.height, c=c)

    def center(self):
        return self.cone2origin[:3, 3] + 0.5 * self.height * self.cone2origin[:3, 2]

    def first_vertex(self):
        return self.cone2origin[:3, 3] + self.height * self.cone2origin[:3, 2]

    def update_pose(self, pose):
        self.cone2origin = pose
        if self.artist_ is not None:
            self.artist_.set_data(pose)

    def collider2origin(self):
        return self.cone2origin

    def pack(self, builder, margin=0.0):
        return builder.add_cone(self.cone2origin, self.radius, self.height, margin=margin)


class Margin(ConvexCollider):
    """Margin around a collider (colliders.py:606-646): support point plus
    margin * d / |d|.  Nested margins add up."""

    def __init__(self, collider, margin):
        super(Margin, self).__init__(collider.artist_)
        self.collider = collider
        self.margin = margin

    def make_artist(self, c=None):
        self.collider.make_artist(c)
        self.artist_ = self.collider.artist_

    def first_vertex(self):
        return self.collider.first_vertex()

    def center(self):
        return self.collider.center()

    def update_pose(self, pose):
        self.collider.update_pose(pose)

    def collider2origin(self):
        return self.collider.collider2origin()

    def pack(self, builder, margin=0.0):
        return self.collider.pack(builder, margin=margin + self.margin)


COLLIDERS = {
    "sphere": Sphere,
    "ellipsoid": Ellipsoid,
    "capsule": Capsule,
    "disk": Disk,
    "ellipse": Ellipse,
    "cone": Cone,
    "cylinder": Cylinder,
    "box": Box,
    "mesh": MeshGraph,
}
