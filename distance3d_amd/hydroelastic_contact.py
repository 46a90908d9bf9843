"""Hydroelastic contact with the reference's call surface
(hydroelastic_contact/_interface.py, _rigid_body.py, _contact_surface.py,
_forces.py, _tetrahedron_intersection.py:7-139, _barycentric_transform.py,
_mesh_processing.py), evaluated by the gfx950 kernels:

- d3d_tet_mesh_prep: per-tetrahedron AABBs and barycentric transforms;
- the LBVH broad phase (all_aabbs_overlap / AabbTree) over those AABBs;
- d3d_tet_pairs_batch: per tetrahedron pair the contact plane where the two
  Young's-modulus-weighted pressure fields agree, and the contact polygon;
- d3d_tet_contact_forces: centre, force and area of each polygon.

The host keeps what the reference keeps in Python objects (the rigid body's
mesh and pose, the contact surface) and sums the per-polygon wrenches.
"""
import numpy as np

from . import kernels
from .aabb_tree import AabbTree
from .broad_phase import BoundingVolumeHierarchy
from .tetra_mesh import (center_of_mass_tetrahedral_mesh, make_tetrahedral_box,
                         make_tetrahedral_capsule, make_tetrahedral_cube,
                         make_tetrahedral_cylinder, make_tetrahedral_ellipsoid,
                         make_tetrahedral_sphere, tetrahedral_mesh_volumes)

tet_pairs_batch = kernels.tet_pairs_batch

SAME = 0x10
CUTS = 0x20
OVERFLOW = 0x100


def _prep(tetrahedra_points, aabbs=True, transforms=True):
    ab, X, flags = kernels.tet_mesh_prep(tetrahedra_points, aabbs, transforms)
    if transforms and len(flags) and bool((flags != 0).any()):
        bad = int(np.nonzero(flags.cpu().numpy())[0][0])
        raise ValueError("tetrahedron %d is degenerate (no barycentric transform)" % bad)
    return ab, X


def barycentric_transforms(tetrahedra_points):
    """X per tetrahedron with X.dot((r, 1)) = barycentric coordinates of r
    (_barycentric_transform.py:4-9), the inverse of [T^T; 1] computed on the
    GPU (the reference's pseudo-inverse is the same matrix for a
    non-degenerate tetrahedron; a degenerate one raises ValueError)."""
    return _prep(np.asarray(tetrahedra_points, dtype=np.float64).reshape(-1, 4, 3),
                 aabbs=False)[1].cpu().numpy()


def tetrahedral_mesh_aabbs(tetrahedra_points):
    """(n, 3, 2) [axis][min, max] box of each tetrahedron
    (_mesh_processing.py:4-20)."""
    return kernels.tet_mesh_prep(tetrahedra_points, True, False)[0].cpu().numpy()


def _dense(X, n):
    """X as an (n, 4, 4) array; the reference passes dicts index -> X."""
    if isinstance(X, dict):
        out = np.zeros((n, 4, 4))
        for k, v in X.items():
            out[k] = v
        return out
    return np.asarray(X, dtype=np.float64).reshape(-1, 4, 4)


def _check(status):
    if np.any(status & OVERFLOW):
        raise AssertionError("more half-plane intersection points than the reference reserves")


def _intersecting(T1, e1, X1, T2, e2, X2, pairs, E1, E2, forces, checked=True):
    """The intersecting pairs of a pair list, in list order, on the GPU: a
    status-only pass over all pairs, the hit (and overflow) rows compacted on
    the device, then planes, polygons (and forces) for the hits only.
    Returns host arrays (hit pairs (k,2), planes (k,4), n_vertices (k,),
    polygons (k,24,3), forces (k,7) or None).  ``checked=False``: the pairs
    are known to be in range (produced by this library's broad phase)."""
    import torch
    pr = torch.as_tensor(pairs)
    pr = (pr if pr.is_cuda else pr.cuda()).to(torch.int32).reshape(-1, 2).contiguous()
    st, _, _, _ = kernels.tet_pairs_batch(T1, e1, X1, T2, e2, X2, pr, E1, E2, geometry=False,
                                          checked=checked)
    rows = torch.nonzero(st & (1 | OVERFLOW)).squeeze(1)
    sel = st[rows]
    _check(sel.cpu().numpy())
    hp = pr[rows[(sel & 1) != 0]]
    if hp.shape[0] == 0:
        return (np.empty((0, 2), dtype=np.int64), np.empty((0, 4)), np.empty(0, dtype=np.int32),
                np.empty((0, kernels.TET_MAX_POLY, 3)), np.empty((0, 7)) if forces else None)
    st2, planes, nv, polys = kernels.tet_pairs_batch(T1, e1, X1, T2, e2, X2, hp, E1, E2,
                                                     checked=False)
    F = (kernels.tet_contact_forces(T1, e1, hp, st2, planes, nv, polys, E1, checked=False)
         if forces else None)
    return (hp.cpu().numpy().astype(np.int64), planes.cpu().numpy(), nv.cpu().numpy(),
            polys.cpu().numpy(), F.cpu().numpy() if forces else None)


def intersect_tetrahedron_pairs(pairs, tetrahedra_points1, tetrahedra_points2, epsilon1, epsilon2,
                                X1, X2, youngs_modulus1=1.0, youngs_modulus2=1.0):
    """Intersect pairs of tetrahedra (_tetrahedron_intersection.py:7-85).

    Returns ``(intersection, contact_planes, contact_polygons,
    intersecting_tetrahedra1, intersecting_tetrahedra2)`` for the
    intersecting pairs, in the order of ``pairs``."""
    T1 = np.asarray(tetrahedra_points1, dtype=np.float64)
    T2 = np.asarray(tetrahedra_points2, dtype=np.float64)
    pr = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if len(pr) == 0:
        return False, [], [], [], []
    # the reference looks X up pair by pair, X1[i] then X2[j]: the first
    # missing entry of a dict in that order is the KeyError it raises
    d1, d2 = isinstance(X1, dict), isinstance(X2, dict)
    if d1 or d2:
        k1 = set(X1) if d1 else None
        k2 = set(X2) if d2 else None
        for i, j in pr.tolist():
            if d1 and i not in k1:
                raise KeyError(i)
            if d2 and j not in k2:
                raise KeyError(j)
    hp, planes, nv, polys, _ = _intersecting(T1, epsilon1, _dense(X1, len(T1)), T2, epsilon2,
                                             _dense(X2, len(T2)), pr, youngs_modulus1,
                                             youngs_modulus2, False)
    return (len(hp) > 0, planes if len(hp) else [], [polys[k, :nv[k]] for k in range(len(hp))],
            hp[:, 0].tolist(), hp[:, 1].tolist())


MAX_HALFPLANES = 1024  # D3D_MAX_HALFPLANES


def intersect_halfplanes(halfplanes):
    """Polygon points of a set of half-planes, rows (p, pq)
    (_halfplanes.py:29-64), on the GPU (d3d_intersect_halfplanes); raises
    AssertionError where the reference does."""
    import torch
    hp = np.ascontiguousarray(halfplanes, dtype=np.float64).reshape(-1, 4)
    n = hp.shape[0]
    if n > MAX_HALFPLANES:
        raise ValueError("at most %d half-planes (one wavefront does O(n^3) work), got %d"
                         % (MAX_HALFPLANES, n))
    hp = torch.as_tensor(hp).cuda()
    pts = torch.empty((max(3 * n, 1), 2), dtype=torch.float64, device="cuda")
    count = torch.empty(1, dtype=torch.int32, device="cuda")
    kernels.check(kernels.lib().d3d_intersect_halfplanes(kernels.ptr(hp), n, kernels.ptr(pts),
                                                         kernels.ptr(count), kernels.stream_ptr(None)),
                  "d3d_intersect_halfplanes")
    k = int(count.item())
    if k < 0:
        raise AssertionError("more half-plane intersection points than the reference reserves")
    return pts[:k].cpu().numpy()


def intersect_tetrahedron_pair(tetrahedron1, epsilon1, X1, tetrahedron2, epsilon2, X2,
                               youngs_modulus1=1.0, youngs_modulus2=1.0):
    """Intersect a pair of tetrahedra (_tetrahedron_intersection.py:89-139):
    ``(intersection, (contact_plane_hnf, contact_polygon))``; the polygon is
    None when the plane does not cut both tetrahedra."""
    st, planes, nv, polys = kernels.tet_pairs_batch(
        np.asarray(tetrahedron1)[None], np.asarray(epsilon1)[None], np.asarray(X1)[None],
        np.asarray(tetrahedron2)[None], np.asarray(epsilon2)[None], np.asarray(X2)[None],
        [[0, 0]], youngs_modulus1, youngs_modulus2)
    st = int(st.cpu().item())
    _check(np.array([st]))
    plane = planes.cpu().numpy()[0]
    n = int(nv.cpu().item())
    if st & 1:
        return True, (plane, polys.cpu().numpy()[0, :n])
    # the plane misses one of the tetrahedra: no polygon; it cuts both but
    # the polygon has fewer than 3 points: an empty one
    return False, (plane, np.empty((0, 3)) if st & CUTS else None)


# ------------------------------------------------------------ contact surface

# the fan of an 8-gon, the reference's TRIANGLES (_forces.py:27-48): polygon
# triangles (0, q, q + 1), at most six
TRIANGLES = np.stack([np.zeros(6, dtype=int), np.arange(1, 7), np.arange(2, 8)], axis=1)


def _invert_transform(A2B):
    """utils.py:211-230."""
    B2A = np.empty((4, 4))
    RT = A2B[:3, :3].T
    B2A[:3, :3] = RT
    B2A[:3, 3] = -np.dot(RT, A2B[:3, 3])
    B2A[3, :3] = 0.0
    B2A[3, 3] = 1.0
    return B2A


def _transform_points(A2B, points):
    """utils.py:148-164."""
    return np.dot(points, A2B[:3, :3].T) + A2B[:3, 3]


def _transform_directions(A2B, directions):
    """utils.py:167-185."""
    return np.dot(directions, A2B[:3, :3].T)


def _adjoint_from_transform(A2B):
    """utils.py:262-324: [[R, 0], [[p]x R, R]]."""
    R, p = A2B[:3, :3], A2B[:3, 3]
    px = np.array([[0.0, -p[2], p[1]], [p[2], 0.0, -p[0]], [-p[1], p[0], 0.0]])
    adj = np.zeros((6, 6))
    adj[:3, :3] = R
    adj[3:, :3] = np.dot(px, R)
    adj[3:, 3:] = R
    return adj


class RigidBody:
    """Rigid body as a tetrahedral mesh (_rigid_body.py:10-296): pose
    body2origin (4, 4), vertices (n, 3), tetrahedra (m, 4), potentials (n,)
    (depth below the surface), Young's modulus."""

    def __init__(self, body2origin, vertices, tetrahedra, potentials, youngs_modulus=1.0):
        self.body2origin_ = np.asarray(body2origin, dtype=np.float64)
        self.vertices_ = np.asarray(vertices, dtype=np.float64)
        self.tetrahedra_ = np.asarray(tetrahedra, dtype=int)
        self.potentials_ = np.asarray(potentials, dtype=np.float64)
        self.youngs_modulus = youngs_modulus
        self._invalidate()

    def _invalidate(self):
        self._tetrahedra_points = None
        self._com = None
        self._aabbs = None
        self._aabb_tree = None

    @staticmethod
    def make_sphere(center, radius, order=4):
        """_rigid_body.py:43-67 (pose = translation to the centre)."""
        mesh2origin = np.eye(4)
        mesh2origin[:3, 3] = center
        return RigidBody(mesh2origin, *make_tetrahedral_sphere(radius, order))

    @staticmethod
    def make_ellipsoid(ellipsoid2origin, radii, order=4):
        """_rigid_body.py:69-91."""
        return RigidBody(ellipsoid2origin, *make_tetrahedral_ellipsoid(radii, order))

    @staticmethod
    def make_cube(cube2origin, size):
        """_rigid_body.py:93-110."""
        return RigidBody(cube2origin, *make_tetrahedral_cube(size))

    @staticmethod
    def make_box(box2origin, size):
        """_rigid_body.py:112-129."""
        return RigidBody(box2origin, *make_tetrahedral_box(size))

    @staticmethod
    def make_cylinder(cylinder2origin, radius, length, resolution_hint=0.1):
        """_rigid_body.py:131-156."""
        return RigidBody(cylinder2origin, *make_tetrahedral_cylinder(radius, length, resolution_hint))

    @staticmethod
    def make_capsule(capsule2origin, radius, height, resolution_hint=0.1):
        """_rigid_body.py:158-184."""
        return RigidBody(capsule2origin, *make_tetrahedral_capsule(radius, height, resolution_hint))

    @property
    def tetrahedra_points(self):
        """(m, 4, 3) points of every tetrahedron in the body frame."""
        if self._tetrahedra_points is None:
            self._tetrahedra_points = self.vertices_[self.tetrahedra_]
        return self._tetrahedra_points

    @property
    def tetrahedra_potentials(self):
        return self.potentials_[self.tetrahedra_]

    @property
    def com(self):
        """Centre of mass at uniform density."""
        if self._com is None:
            self._com = center_of_mass_tetrahedral_mesh(self.tetrahedra_points)
        return self._com

    def express_in(self, new_body2origin):
        """Vertices re-expressed in another frame (_rigid_body.py:208-226)."""
        body2new = np.dot(_invert_transform(new_body2origin), self.body2origin_)
        self.vertices_ = _transform_points(body2new, self.vertices_)
        self.body2origin_ = np.copy(new_body2origin)
        self._invalidate()

    def update_pose(self, body2origin):
        """_rigid_body.py:228-239 (the pose only; vertices stay where
        express_in put them, as in the reference)."""
        self.body2origin_ = body2origin

    @property
    def artist_(self):
        """No visualisation here (the reference's wireframe artist)."""
        return None

    def aabb(self):
        """Box around the whole body."""
        return self.aabb_tree.get_root_aabb()

    @property
    def aabbs(self):
        """(m, 3, 2) tetrahedron boxes for the broad phase."""
        if self._aabbs is None:
            self._aabbs = tetrahedral_mesh_aabbs(self.tetrahedra_points)
        return self._aabbs

    @property
    def aabb_tree(self):
        if self._aabb_tree is None:
            self._aabb_tree = AabbTree()
            self._aabb_tree.insert_aabbs(self.aabbs, pre_insertion_methode="sort")
        return self._aabb_tree


class ContactSurface:
    """Contact surface of two tetrahedral meshes (_contact_surface.py:5-141):
    per intersecting tetrahedron pair its plane (Hesse normal form), polygon,
    and after add_polygon_info its area, centre, force and triangles."""

    def __init__(self, frame2world, intersection, contact_planes, contact_polygons,
                 intersecting_tetrahedra1, intersecting_tetrahedra2):
        self.frame2world = frame2world
        self.intersection = intersection
        self.contact_planes = contact_planes
        self.contact_polygons = contact_polygons
        self.intersecting_tetrahedra1 = intersecting_tetrahedra1
        self.intersecting_tetrahedra2 = intersecting_tetrahedra2
        self.contact_areas = None
        self.contact_coms = None
        self.contact_forces = None
        self.contact_polygon_triangles = None

    def add_polygon_info(self, contact_areas, contact_coms, contact_forces,
                         contact_polygon_triangles):
        self.contact_areas = contact_areas
        self.contact_coms = contact_coms
        self.contact_forces = contact_forces
        self.contact_polygon_triangles = contact_polygon_triangles

    @property
    def pressures(self):
        """|force| / area per polygon (zero areas counted as 1)."""
        areas = np.copy(self.contact_areas)
        areas[areas == 0.0] = 1.0
        return np.linalg.norm(self.contact_forces, axis=1) / areas

    def make_details(self, tetrahedra_points1, tetrahedra_points2):
        """Contact details in the world frame (_contact_surface.py:79-123)."""
        if not self.intersection:
            return {}
        t1, t2 = self._transform_to_world(self.frame2world, tetrahedra_points1, tetrahedra_points2)
        contact_point = np.sum(self.contact_coms * self.contact_areas[:, np.newaxis],
                               axis=0) / sum(self.contact_areas)
        return {
            "contact_polygons": self.contact_polygons,
            "contact_polygon_triangles": self.contact_polygon_triangles,
            "contact_planes": self.contact_planes,
            "intersecting_tetrahedra1": t1,
            "intersecting_tetrahedra2": t2,
            "contact_coms": self.contact_coms,
            "contact_forces": self.contact_forces,
            "contact_areas": self.contact_areas,
            "pressures": self.pressures,
            "contact_point": contact_point,
        }

    def _transform_to_world(self, frame2world, tetrahedra_points1, tetrahedra_points2):
        """_contact_surface.py:125-141."""
        self.contact_polygons = [_transform_points(frame2world, p) for p in self.contact_polygons]
        plane_points = _transform_points(
            frame2world, self.contact_planes[:, :3] * self.contact_planes[:, 3, np.newaxis])
        plane_normals = _transform_directions(frame2world, self.contact_planes[:, :3])
        plane_distances = np.sum(plane_points * plane_normals, axis=1)
        self.contact_planes = np.hstack((plane_normals, plane_distances.reshape(-1, 1)))
        self.contact_coms = _transform_points(frame2world, self.contact_coms)
        self.contact_forces = _transform_directions(frame2world, self.contact_forces)
        n = len(self.intersecting_tetrahedra1)
        t1 = np.asarray(tetrahedra_points1)[np.asarray(self.intersecting_tetrahedra1, dtype=int)]
        t2 = np.asarray(tetrahedra_points2)[np.asarray(self.intersecting_tetrahedra2, dtype=int)]
        return (_transform_points(frame2world, t1.reshape(n * 4, 3)).reshape(n, 4, 3),
                _transform_points(frame2world, t2.reshape(n * 4, 3)).reshape(n, 4, 3))


def find_contact_surface(rigid_body1, rigid_body2, use_aabb_trees=False):
    """Contact surface of two rigid bodies (_interface.py:52-101); like the
    reference it re-expresses rigid_body1 in rigid_body2's frame.  Broad phase
    over the tetrahedron boxes, then every overlapping tetrahedron pair on
    the GPU."""
    import torch
    rigid_body1.express_in(rigid_body2.body2origin_)
    T1 = torch.from_numpy(np.ascontiguousarray(rigid_body1.tetrahedra_points)).cuda()
    T2 = torch.from_numpy(np.ascontiguousarray(rigid_body2.tetrahedra_points)).cuda()
    ab1, X1 = _prep(T1)
    ab2, X2 = _prep(T2)
    if use_aabb_trees:
        _, _, _, pairs = rigid_body1.aabb_tree.overlaps_aabb_tree(rigid_body2.aabb_tree)
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    elif len(ab1) * len(ab2) <= kernels.DENSE_OVERLAP_MAX_TESTS:
        pairs = kernels.overlap_dense(ab1, ab2)
    elif len(ab1) and len(ab2):
        pairs = kernels.bvh_query(kernels.bvh_build(ab2), ab1)
    else:
        pairs = np.empty((0, 2), dtype=np.int64)
    E1, E2 = rigid_body1.youngs_modulus, rigid_body2.youngs_modulus
    if len(pairs) == 0:
        surface = ContactSurface(rigid_body2.body2origin_, False, np.empty((0, 4)), [], [], [])
        surface.add_polygon_info(np.empty(0), np.empty((0, 3)), np.empty((0, 3)), [])
        return surface
    hp, planes, nv, polys, F = _intersecting(
        T1, rigid_body1.tetrahedra_potentials, X1, T2, rigid_body2.tetrahedra_potentials, X2,
        pairs, E1, E2, True, checked=use_aabb_trees)
    k = len(hp)
    surface = ContactSurface(rigid_body2.body2origin_, k > 0, planes,
                             [polys[q, :nv[q]] for q in range(k)], hp[:, 0].tolist(),
                             hp[:, 1].tolist())
    surface.add_polygon_info(F[:, 6], F[:, 0:3], F[:, 3:6], [TRIANGLES[:nv[q] - 2] for q in range(k)])
    return surface


def contact_surface_forces(contact_surface, rigid_body1):
    """(areas, centres, forces, triangles) of a contact surface's polygons
    (_forces.py:6-24), on the GPU."""
    n = len(contact_surface.intersecting_tetrahedra1)
    if n == 0:
        return np.empty(0), np.empty((0, 3)), np.empty((0, 3)), []
    polys = np.full((n, kernels.TET_MAX_POLY, 3), np.nan)
    nv = np.array([len(p) for p in contact_surface.contact_polygons], dtype=np.int32)
    for k, p in enumerate(contact_surface.contact_polygons):
        polys[k, :len(p)] = p
    pairs = np.stack([np.asarray(contact_surface.intersecting_tetrahedra1), np.zeros(n, dtype=int)], 1)
    out = kernels.tet_contact_forces(
        rigid_body1.tetrahedra_points, rigid_body1.tetrahedra_potentials, pairs,
        np.ones(n, dtype=np.int32), np.asarray(contact_surface.contact_planes), nv, polys,
        rigid_body1.youngs_modulus).cpu().numpy()
    return out[:, 6], out[:, 0:3], out[:, 3:6], [TRIANGLES[:k - 2] for k in nv]


def compute_contact_force(tetrahedron, epsilon, contact_plane_hnf, contact_polygon,
                          youngs_modulus=1.0):
    """One polygon's (centre, force vector, area, triangles)
    (_forces.py:52-111)."""
    poly = np.asarray(contact_polygon, dtype=np.float64).reshape(-1, 3)
    if len(poly) > kernels.TET_MAX_POLY:
        raise ValueError("contact polygon has more than %d vertices" % kernels.TET_MAX_POLY)
    padded = np.full((1, kernels.TET_MAX_POLY, 3), np.nan)
    padded[0, :len(poly)] = poly
    out = kernels.tet_contact_forces(
        np.asarray(tetrahedron, dtype=np.float64).reshape(1, 4, 3),
        np.asarray(epsilon, dtype=np.float64).reshape(1, 4), [[0, 0]], [1],
        np.asarray(contact_plane_hnf, dtype=np.float64).reshape(1, 4), [len(poly)], padded,
        youngs_modulus).cpu().numpy()[0]
    return out[0:3], out[3:6], out[6], TRIANGLES[:max(len(poly) - 2, 0)]


def accumulate_wrenches(contact_surface, rigid_body1, rigid_body2):
    """Net wrenches of the polygon forces about each body's centre of mass,
    in the world frame (_forces.py:114-139)."""
    f = contact_surface.contact_forces
    c = contact_surface.contact_coms
    force21 = np.sum(f, axis=0)
    torque21 = np.sum(np.cross(c - rigid_body1.com, f), axis=0)
    torque12 = np.sum(np.cross(c - rigid_body2.com, -f), axis=0)
    adj = _adjoint_from_transform(contact_surface.frame2world)
    wrench21 = adj.T.dot(np.hstack((force21, torque21)))
    wrench12 = adj.T.dot(np.hstack((-force21, torque12)))
    return wrench12, wrench21


def contact_forces(rigid_body1, rigid_body2, return_details=False):
    """(intersection, wrench12_in_world, wrench21_in_world[, details]) of two
    rigid bodies (_interface.py:8-49)."""
    surface = find_contact_surface(rigid_body1, rigid_body2)
    wrench12, wrench21 = accumulate_wrenches(surface, rigid_body1, rigid_body2)
    if return_details:
        details = surface.make_details(rigid_body1.tetrahedra_points, rigid_body2.tetrahedra_points)
        return surface.intersection, wrench12, wrench21, details
    return surface.intersection, wrench12, wrench21


class HydroelasticBoundingVolumeHierarchy(BoundingVolumeHierarchy):
    """Broad phase over RigidBody colliders (_broad_phase.py:7-98): the
    BoundingVolumeHierarchy of this package with each body's box from its
    tetrahedron boxes (RigidBody.aabb).  Bodies are added with add_collider;
    filling from a URDF needs pytransform3d, which is not available."""

    def aabb_overlapping_colliders(self, collider, whitelist=()):
        """{frame: body} of the bodies whose box overlaps ``collider``'s."""
        return super().aabb_overlapping_colliders(collider, whitelist)
