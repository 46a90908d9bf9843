"""Tetrahedral meshes of primitive shapes and their mass properties, in the
reference's data format (hydroelastic_contact/_tetra_mesh_creation.py,
_mesh_processing.py): vertices (n_vertices, 3), tetrahedra (n_tetrahedra, 4)
vertex indices, potentials (n_vertices,) = depth below the surface.  The same
vertex and tetrahedron order as the reference, so meshes index alike.

Host-side mesh construction (once per body); the per-contact work runs on the
GPU (hydroelastic_contact).
"""
import numpy as np

# the twelve icosahedron vertices and twenty faces of the icosphere seed
# (_tetra_mesh_creation.py:31-42, after mourner's fast-icosphere-mesh)
_PHI = (1 + 5 ** 0.5) / 2
_ICO_VERTICES = np.array([
    [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0], [0, -1, _PHI], [0, 1, _PHI],
    [0, -1, -_PHI], [0, 1, -_PHI], [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]])
_ICO_FACES = np.array([
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [11, 10, 2], [5, 11, 4],
    [1, 5, 9], [7, 1, 8], [10, 7, 6], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [9, 8, 1], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7]], dtype=int)


def make_triangular_icosphere(center, radius, order=4):
    """Icosphere surface (_tetra_mesh_creation.py:6-80): each subdivision
    splits a triangle (v1, v2, v3) into (v1, a, c), (v2, b, a), (v3, c, b),
    (a, b, c) with edge midpoints a, b, c numbered in order of first use;
    the vertices are then scaled to the radius and moved to the centre."""
    vertices = np.zeros((10 * 4 ** order + 2, 3))
    vertices[:12] = _ICO_VERTICES
    triangles = _ICO_FACES
    n = 12
    midpoint = {}
    for _ in range(order):
        refined = np.empty((4 * len(triangles), 3), dtype=int)
        for k, (v1, v2, v3) in enumerate(triangles):
            mids = []
            for a, b in ((v1, v2), (v2, v3), (v3, v1)):
                key = (min(a, b), max(a, b))
                idx = midpoint.pop(key, None)  # an edge is shared by two faces
                if idx is None:
                    idx = midpoint[key] = n
                    vertices[n] = 0.5 * (vertices[a] + vertices[b])
                    n += 1
                mids.append(idx)
            a, b, c = mids
            refined[4 * k:4 * k + 4] = ((v1, a, c), (v2, b, a), (v3, c, b), (a, b, c))
        triangles = refined
    vertices /= 1.0 / radius * np.linalg.norm(vertices, axis=1)[:, np.newaxis]
    vertices += np.asarray(center, dtype=np.float64)[np.newaxis]
    return vertices, triangles


def _fan_to_centre(vertices, triangles, centre_potential):
    """Tetrahedra (triangle, centre vertex) with the centre appended last
    and given the potential (_tetra_mesh_creation.py:105-112)."""
    centre = len(vertices)
    vertices = np.vstack((vertices, np.zeros((1, 3))))
    tetrahedra = np.hstack((triangles, np.full((len(triangles), 1), centre, dtype=int)))
    potentials = np.zeros(len(vertices))
    potentials[-1] = centre_potential
    return vertices, tetrahedra, potentials


def make_tetrahedral_sphere(radius, order=4):
    """_tetra_mesh_creation.py:83-112."""
    vertices, triangles = make_triangular_icosphere(np.zeros(3), radius, order)
    return _fan_to_centre(vertices, triangles, radius)


def make_tetrahedral_ellipsoid(radii, order=4):
    """_tetra_mesh_creation.py:115-145."""
    radii = np.asarray(radii, dtype=np.float64)
    vertices, triangles = make_triangular_icosphere(np.zeros(3), 1.0, order)
    vertices *= radii[np.newaxis]
    return _fan_to_centre(vertices, triangles, min(radii))


_CUBE_TETRAHEDRA = np.array([
    [0, 2, 6, 8], [0, 4, 5, 8], [0, 1, 2, 8], [1, 3, 2, 8], [1, 5, 7, 8], [1, 7, 3, 8],
    [5, 1, 0, 8], [5, 6, 7, 8], [6, 2, 3, 8], [6, 3, 7, 8], [6, 4, 0, 8], [6, 5, 4, 8]], dtype=int)


def make_tetrahedral_cube(size):
    """_tetra_mesh_creation.py:148-195: the 8 corners in binary order
    (x slowest) plus the centre, 12 tetrahedra (two per face)."""
    corners = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)]
                       + [[0.0, 0.0, 0.0]])
    vertices = size * corners
    potentials = np.zeros(9)
    potentials[8] = size / 2.0
    return vertices, _CUBE_TETRAHEDRA.copy(), potentials


def _split_to_tetrahedra(v0, v1, v2, v3, v4, v5, v6, v7):
    """Tetrahedra (prev, next, v0, v6) around the diagonal v0-v6 of a
    hexahedron, skipping degenerate ones (_tetra_mesh_creation.py:292-299)."""
    out = []
    prev = v1
    for nxt in (v2, v3, v7, v4, v5, v1):
        if len({prev, nxt, v0, v6}) == 4:
            out.append([prev, nxt, v0, v6])
        prev = nxt
    return out


def make_tetrahedral_box(size):
    """_tetra_mesh_creation.py:198-289 (Drake's box mesh): the 8 corners,
    then the corners of the medial box (half sizes minus the smallest; a
    collapsed axis shares vertices), and six hexahedra between each face and
    the medial box split around their diagonals."""
    half = 0.5 * np.asarray(size, dtype=np.float64)
    verts = []
    v = np.empty((2, 2, 2), dtype=int)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                v[i, j, k] = len(verts)
                verts.append([half[0] if i else -half[0], half[1] if j else -half[1],
                              half[2] if k else -half[2]])
    min_half = min(half)
    tol = 1e-14 * max(1.0, min_half)
    central = half - min_half
    central[central <= tol] = 0.0
    m = np.empty((2, 2, 2), dtype=int)
    for i in range(2):
        for j in range(2):
            for k in range(2):
                dup_i = i == 1 and central[0] == 0.0
                dup_j = j == 1 and central[1] == 0.0
                dup_k = k == 1 and central[2] == 0.0
                if dup_i:
                    m[i, j, k] = m[0, j, k]
                elif dup_j:
                    m[i, j, k] = m[i, 0, k]
                elif dup_k:
                    m[i, j, k] = m[i, j, 0]
                else:
                    m[i, j, k] = len(verts)
                    verts.append([central[0] if i else -central[0], central[1] if j else -central[1],
                                  central[2] if k else -central[2]])
    verts = np.array(verts, dtype=np.float64)
    faces = (((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1)),
             ((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)),
             ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),
             ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)),
             ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
             ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)))
    tets = []
    for face in faces:
        tets.extend(_split_to_tetrahedra(*[int(m[c]) for c in face], *[int(v[c]) for c in face]))
    potentials = np.zeros(len(verts))
    potentials[8:] = min_half
    return verts, np.array(tets, dtype=int), potentials


def _split_prism(v0, v1, v2, v3, v4, v5):
    """Three tetrahedra of the triangular prism (v0 v1 v2 | v3 v4 v5) around
    the diagonal v0-v5 (_tetra_mesh_creation.py:468-474)."""
    out, prev = [], v3
    for nxt in (v4, v1, v2):
        out.append([prev, nxt, v0, v5])
        prev = nxt
    return out


def _split_pyramid(v0, v1, v2, v3, apex):
    """Two tetrahedra of the pyramid over the quad (v0 v1 v2 v3)
    (_tetra_mesh_creation.py:477-483)."""
    return [[v3, apex, v0, v2], [apex, v1, v0, v2]]


def make_tetrahedral_cylinder(radius, length, resolution_hint):
    """Cylinder along z (_tetra_mesh_creation.py:302-465, Drake's cylinder
    mesh): the two cap centres, then bottom/top rim vertices interleaved
    (ceil(2 pi r / hint) per circle, at least 3), then the medial axis:
    a segment (long: half length > r), a point (medium) or a circle of
    radius r - L/2 at z = 0 plus the centre (short)."""
    import math
    top_z = 0.5 * length
    tol = 1e-14 * max(1.0, min(top_z, radius))
    n = max(3, math.ceil(2.0 * np.pi * radius / resolution_hint))
    verts = [np.array([0.0, 0.0, -top_z]), np.array([0.0, 0.0, top_z])]
    bc, tc = 0, 1
    bottom, top = [], []
    step = 2.0 * np.pi / n
    for i in range(n):
        x, y = radius * np.cos(step * i), radius * np.sin(step * i)
        bottom.append(len(verts))
        verts.append(np.array([x, y, -top_z]))
        top.append(len(verts))
        verts.append(np.array([x, y, top_z]))
    pot = [0.0] * len(verts)
    tets = []
    if top_z - radius > tol:  # long: medial segment
        m0, m1 = len(verts), len(verts) + 1
        off = top_z - radius
        verts += [np.array([0.0, 0.0, -off]), np.array([0.0, 0.0, off])]
        pot += [radius, radius]
        i = n - 1
        for j in range(n):
            tets.append([bc, bottom[i], bottom[j], m0])
            tets.append([tc, top[j], top[i], m1])
            tets.extend(_split_prism(m0, bottom[i], bottom[j], m1, top[i], top[j]))
            i = j
    elif radius - top_z > tol:  # short: medial circle and centre
        centre = len(verts)
        verts.append(np.array([0.0, 0.0, 0.0]))
        pot.append(top_z)
        scale = (radius - top_z) / radius
        medial = []
        for i in range(n):
            medial.append(len(verts))
            verts.append(np.array([verts[bottom[i]][0] * scale, verts[bottom[i]][1] * scale, 0.0]))
            pot.append(top_z)
        i = n - 1
        for j in range(n):
            tets.extend(_split_prism(bc, bottom[i], bottom[j], centre, medial[i], medial[j]))
            tets.extend(_split_prism(centre, medial[i], medial[j], tc, top[i], top[j]))
            tets.extend(_split_prism(bottom[i], medial[i], top[i], bottom[j], medial[j], top[j]))
            i = j
    else:  # medium: medial point at the centre
        m = len(verts)
        verts.append(np.array([0.0, 0.0, 0.0]))
        pot.append(radius)
        i = n - 1
        for j in range(n):
            tets.append([bc, bottom[i], bottom[j], m])
            tets.append([tc, top[j], top[i], m])
            tets.extend(_split_pyramid(top[i], top[j], bottom[j], bottom[i], m))
            i = j
    return np.array(verts), np.array(tets, dtype=int), np.array(pot)


def make_tetrahedral_capsule(radius, height, resolution_hint):
    """Capsule along z (_tetra_mesh_creation.py:486-593, Drake's capsule
    mesh): medial top/bottom, the two poles, then the cap circles (top and
    bottom interleaved, n = clip(2 pi r / hint, 3, 706) vertices each, n // 2
    circles per cap from the equator up); pyramids between circles, fans at
    the poles, prisms along the side around the medial segment."""
    mz = 0.5 * height
    pz = mz + radius
    n = int(np.clip(2.0 * np.pi * radius / resolution_hint, 3.0, 706.0))
    nc = n // 2
    verts = [np.array([0.0, 0.0, mz]), np.array([0.0, 0.0, -mz]),
             np.array([0.0, 0.0, pz]), np.array([0.0, 0.0, -pz])]
    mtop, mbot, ptop, pbot = 0, 1, 2, 3
    tcap, bcap = [], []
    th_step = 0.5 * np.pi / nc
    ph_step = 2.0 * np.pi / n
    for i in range(nc):
        theta = 0.5 * np.pi - i * th_step
        st = np.sin(theta)
        z = radius * np.cos(theta) + mz
        for j in range(n):
            phi = j * ph_step
            x, y = radius * st * np.cos(phi), radius * st * np.sin(phi)
            tcap.append(len(verts))
            verts.append(np.array([x, y, z]))
            bcap.append(len(verts))
            verts.append(np.array([x, y, -z]))
    tets = []
    for i in range(nc - 1):
        for j in range(n):
            j1 = (j + 1) % n
            tets.extend(_split_pyramid(tcap[(i + 1) * n + j], tcap[(i + 1) * n + j1],
                                       tcap[i * n + j1], tcap[i * n + j], mtop))
            tets.extend(_split_pyramid(bcap[i * n + j], bcap[i * n + j1],
                                       bcap[(i + 1) * n + j1], bcap[(i + 1) * n + j], mbot))
    last = (nc - 1) * n
    for j in range(n):
        j1 = (j + 1) % n
        tets.append([ptop, tcap[last + j1], tcap[last + j], mtop])
        tets.append([pbot, bcap[last + j], bcap[last + j1], mbot])
        tets.extend(_split_prism(mbot, bcap[j], bcap[j1], mtop, tcap[j], tcap[j1]))
    pot = np.zeros(len(verts))
    pot[:2] = radius
    return np.array(verts), np.array(tets, dtype=int), pot


def tetrahedral_mesh_volumes(tetrahedra_points):
    """|((p1 - p0) x (p2 - p0)) . (p3 - p0)| / 6 per tetrahedron
    (_mesh_processing.py:43-60)."""
    P = np.asarray(tetrahedra_points, dtype=np.float64)
    e = P[:, 1:] - P[:, np.newaxis, 0]
    return np.abs(np.sum(np.cross(e[:, 0], e[:, 1]) * e[:, 2], axis=1)) / 6.0


def center_of_mass_tetrahedral_mesh(tetrahedra_points):
    """Volume-weighted mean of the tetrahedron centroids
    (_mesh_processing.py:23-40)."""
    P = np.asarray(tetrahedra_points, dtype=np.float64)
    volumes = tetrahedral_mesh_volumes(P)
    return np.dot(volumes, P.mean(axis=1)) / np.sum(volumes)
