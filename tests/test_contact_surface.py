"""Hydroelastic contact surface and forces (hydroelastic_contact/
_interface.py:8-101, _forces.py, _mesh_processing.py, _rigid_body.py,
_tetra_mesh_creation.py).

The oracle is pinned by the reference's own golden wrenches
(test_hydroelastic_contact.py:7-37, 256-293), by a numpy restatement of
compute_contact_force that uses np.linalg.solve / np.linalg.norm as the
reference does, by np.linalg.pinv for the barycentric transforms, and by
what the geometry fixes (polygon area by the shoelace formula; for a linear
pressure field the triangle-centroid quadrature is exact, so the force is the
pressure at the polygon centre times its area).  On the GPU the kernels and
the drop-in pipeline are compared with the oracle bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle
from distance3d_amd import hydroelastic_contact as hc
from distance3d_amd import tetra_mesh


def random_tets(rs, n, spread=0.6):
    return rs.uniform(-spread, spread, (n, 1, 3)) + rs.randn(n, 4, 3) * 0.5


def homogeneous(T):
    T = np.asarray(T, dtype=np.float64).reshape(-1, 4, 3)
    return np.concatenate([T.transpose(0, 2, 1), np.ones((len(T), 1, 4))], axis=1)


def numpy_contact_force(tet, eps, plane, poly, E=1.0):
    """compute_contact_force (_forces.py:52-111) in numpy, solve and norm
    from numpy.linalg as the reference calls them."""
    X = np.vstack((tet.T, np.ones((1, 4))))
    tf, ta, com = 0.0, 0.0, np.zeros(3)
    c4 = np.ones(4)
    for q in range(1, min(len(poly) - 2, 6) + 1):
        v = poly[[0, q, q + 1]]
        c4[:3] = (v[0] + v[1] + v[2]) / 3.0
        pressure = np.sum(np.linalg.solve(X, c4) * (eps * E))
        area = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
        tf += pressure * area
        ta += area
        com += area * c4[:3]
    com = com / ta if ta > 0.0 else poly[0]
    return np.concatenate([com, tf * plane[:3], [ta]])


def overlapping_pairs(ab1, ab2):
    cap = max(len(ab1) * len(ab2), 1)
    pairs = np.empty((cap, 2), dtype=np.int64)
    k = oracle.lib().orc_all_aabbs_overlap(oracle._p(np.ascontiguousarray(ab1)), len(ab1),
                                           oracle._p(np.ascontiguousarray(ab2)), len(ab2),
                                           oracle._p(pairs), cap)
    return pairs[:k]


def oracle_contact_surface(rb1, rb2):
    """find_contact_surface (_interface.py:52-101) from oracle pieces only:
    brute-force box overlaps, then the oracle's pair contact and forces."""
    rb1.express_in(rb2.body2origin_)
    T1, T2 = rb1.tetrahedra_points, rb2.tetrahedra_points
    ab1, X1, _ = oracle.tet_mesh_prep(T1)
    ab2, X2, _ = oracle.tet_mesh_prep(T2)
    pairs = overlapping_pairs(ab1, ab2)
    st, pl, nv, po = oracle.tet_pairs(T1, rb1.tetrahedra_potentials, X1, T2,
                                      rb2.tetrahedra_potentials, X2, pairs,
                                      rb1.youngs_modulus, rb2.youngs_modulus)
    F = oracle.tet_contact_forces(T1, rb1.tetrahedra_potentials, pairs, st, pl, nv, po,
                                  rb1.youngs_modulus)
    hit = np.nonzero(st & 1)[0]
    s = hc.ContactSurface(rb2.body2origin_, len(hit) > 0, pl[hit], [po[q, :nv[q]] for q in hit],
                          pairs[hit, 0].tolist(), pairs[hit, 1].tolist())
    s.add_polygon_info(F[hit, 6], F[hit, :3], F[hit, 3:6], [hc.TRIANGLES[:nv[q] - 2] for q in hit])
    return s


def oracle_contact_forces(rb1, rb2):
    s = oracle_contact_surface(rb1, rb2)
    return (s.intersection,) + hc.accumulate_wrenches(s, rb1, rb2)


def rotation(rs):
    q = rs.randn(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def pose(rs, t):
    P = np.eye(4)
    P[:3, :3] = rotation(rs)
    P[:3, 3] = t
    return P


# ------------------------------------------------------------------ CPU pins

def test_oracle_mesh_prep_vs_numpy():
    """AABBs equal np.min / np.max exactly; the LU inverse equals
    np.linalg.pinv to rounding and inverts [T^T; 1]."""
    rs = np.random.RandomState(0)
    T = random_tets(rs, 500)
    ab, X, flags = oracle.tet_mesh_prep(T)
    assert not flags.any()
    assert np.array_equal(ab[:, :, 0], T.min(axis=1)) and np.array_equal(ab[:, :, 1], T.max(axis=1))
    A = homogeneous(T)
    P = np.linalg.pinv(A)
    scale = np.abs(P).max(axis=(1, 2), keepdims=True)
    assert np.all(np.abs(X - P) <= 1e-11 * scale)
    assert np.allclose(X @ A, np.eye(4), rtol=0, atol=1e-11)
    # a flat tetrahedron has a zero pivot
    flat = T[:1].copy()
    flat[0, 3] = (flat[0, 0] + flat[0, 1]) / 2.0
    flat[0, 2] = flat[0, 0]
    assert oracle.tet_mesh_prep(flat)[2][0] == 1


def test_oracle_forces_vs_numpy_and_geometry():
    rs = np.random.RandomState(1)
    n = 1500
    T1, T2 = random_tets(rs, n), random_tets(rs, n)
    e1, e2 = rs.uniform(0.0, 1.0, (n, 4)), rs.uniform(0.0, 1.0, (n, 4))
    _, X1, _ = oracle.tet_mesh_prep(T1)
    _, X2, _ = oracle.tet_mesh_prep(T2)
    pairs = np.stack([np.arange(n), np.arange(n)], 1)
    E1 = 2.5
    st, pl, nv, po = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs, E1, 0.7)
    F = oracle.tet_contact_forces(T1, e1, pairs, st, pl, nv, po, E1)
    hit = np.nonzero(st & 1)[0]
    assert len(hit) > 30
    assert np.all(F[st & 1 == 0] == 0.0)
    for k in hit:
        poly = po[k, :nv[k]]
        want = numpy_contact_force(T1[k], e1[k], pl[k], poly, E1)
        assert np.allclose(F[k], want, rtol=1e-12, atol=1e-15)
        # shoelace area of the (planar, convex) polygon
        nrm = pl[k, :3]
        shoelace = 0.5 * abs(sum(np.dot(nrm, np.cross(poly[q], poly[(q + 1) % len(poly)]))
                                 for q in range(len(poly))))
        if len(poly) <= 8:
            assert F[k, 6] == pytest.approx(shoelace, rel=1e-10, abs=1e-14)
        # linear pressure field: force = p(centre) * area along the normal
        bc = np.linalg.solve(homogeneous(T1[k])[0], np.append(F[k, :3], 1.0))
        p_centre = np.dot(bc, e1[k] * E1)
        assert np.allclose(F[k, 3:6], p_centre * F[k, 6] * nrm, rtol=1e-9, atol=1e-14)


def test_oracle_same_tetrahedron_has_no_area():  # test_hydroelastic_contact.py:101-116
    T = tetra_mesh.make_tetrahedral_sphere(1.0, 2)
    rb = hc.RigidBody(np.eye(4), *T)
    rb.express_in(np.eye(4))
    tet = rb.tetrahedra_points[:1]
    eps = np.array([[0.0, 0.0, 0.0, 1.0]])
    _, X, _ = oracle.tet_mesh_prep(tet)
    st, pl, nv, po = oracle.tet_pairs(tet, eps, X, tet, eps, X, [[0, 0]])
    F = oracle.tet_contact_forces(tet, eps, [[0, 0]], st, pl, nv, po)
    assert st[0] & 1 and F[0, 6] == 0.0 and np.all(F[0, 3:6] == 0.0)


def test_oracle_golden_wrenches():
    """test_hydroelastic_contact.py:7-37 and 256-293 through the oracle."""
    rb1 = hc.RigidBody.make_sphere(np.array([0.0, 0.0, 0.01]), 0.15, 1)
    for z, fz in ((0.4, None), (0.3, 1.120994e-06), (0.15, 0.00166866)):
        rb2 = hc.RigidBody.make_sphere(np.array([0.0, 0.0, z]), 0.15, 1)
        hit, w12, w21 = oracle_contact_forces(rb1, rb2)
        if fz is None:
            assert not hit and np.all(w12 == 0.0) and np.all(w21 == 0.0)
            continue
        assert hit
        np.testing.assert_array_almost_equal(w12, [0.0, 0.0, fz, 0.0, 0.0, 0.0])
        np.testing.assert_array_almost_equal(w21, [0.0, 0.0, -fz, 0.0, 0.0, 0.0])
    rb1 = hc.RigidBody.make_sphere(np.array([0.5, 0, 0]), 0.15, 1)
    rb2 = hc.RigidBody.make_sphere(np.array([0.5, 0.12, 0]), 0.15, 1)
    _, w12_1, _ = oracle_contact_forces(rb1, rb2)
    rb1.youngs_modulus, rb2.youngs_modulus = 10, 0.1
    hit, w12_3, _ = oracle_contact_forces(rb1, rb2)
    assert hit
    np.testing.assert_array_almost_equal(w12_3, [0, 0.00032, 0, 0, 0, 0], decimal=4)
    assert np.linalg.norm(w12_3) < np.linalg.norm(w12_1)


def test_mesh_mass_properties():
    """test_hydroelastic_contact.py:139-253 for the sphere, ellipsoid, cube
    and box meshes (centre of mass, total volume)."""
    center = np.array([0.0, 0.2, -0.3])
    rb = hc.RigidBody.make_sphere(center, 0.2)
    rb.express_in(np.eye(4))
    np.testing.assert_array_almost_equal(rb.com, center)
    radii = np.array([1.0, 2.0, 3.0])
    for rb, vol, tol in ((hc.RigidBody.make_ellipsoid(np.eye(4), radii), None, None),
                         (hc.RigidBody.make_cube(np.eye(4), 1.0), 1.0, 1e-12),
                         (hc.RigidBody.make_box(np.eye(4), radii), 6.0, 1e-12)):
        np.testing.assert_array_almost_equal(rb.com, np.zeros(3))
        if vol is not None:
            assert np.sum(tetra_mesh.tetrahedral_mesh_volumes(rb.tetrahedra_points)) == pytest.approx(vol, rel=tol)
    sphere = hc.RigidBody.make_sphere(np.array([1.0, 2.0, 3.0]), 1.0)
    sphere.express_in(np.eye(4))
    assert np.sum(tetra_mesh.tetrahedral_mesh_volumes(sphere.tetrahedra_points)) == pytest.approx(
        4.0 / 3.0 * np.pi, abs=1e-2)
    ell = hc.RigidBody.make_ellipsoid(np.eye(4), radii, 5)
    assert np.sum(tetra_mesh.tetrahedral_mesh_volumes(ell.tetrahedra_points)) == pytest.approx(
        4.0 / 3.0 * np.pi * 6.0, abs=1e-1)


def test_mesh_topology():
    """Icosphere counts (10 * 4^order + 2 vertices, 20 * 4^order faces,
    every edge shared by two faces); box meshes tile the box (volume) with
    the medial-axis potentials."""
    for order in range(4):
        v, tri = tetra_mesh.make_triangular_icosphere(np.zeros(3), 2.0, order)
        assert len(v) == 10 * 4 ** order + 2 and len(tri) == 20 * 4 ** order
        assert np.allclose(np.linalg.norm(v, axis=1), 2.0)
        edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        _, counts = np.unique(edges, axis=0, return_counts=True)
        assert np.all(counts == 2)
    for size in ([1.0, 1.0, 1.0], [1.0, 2.0, 3.0], [2.0, 2.0, 0.5]):
        v, t, p = tetra_mesh.make_tetrahedral_box(np.array(size))
        assert np.sum(tetra_mesh.tetrahedral_mesh_volumes(v[t])) == pytest.approx(np.prod(size))
        assert np.all(p[:8] == 0.0) and np.all(p[8:] == min(size) / 2)


def test_cylinder_and_capsule_meshes():
    """Long, medium and short cylinders fill the inscribed polygonal prism
    exactly (volume n/2 r^2 sin(2 pi/n) L, no degenerate tetrahedra); the
    capsule approaches the round capsule's volume; centres of mass at the
    origin (test_hydroelastic_contact.py:139-253)."""
    import math
    for r, L, hint in ((0.1, 1.0, 0.1), (0.1, 0.2, 0.1), (0.1, 0.1, 0.1), (0.1, 1.0, 0.01),
                       (0.3, 0.2, 0.05), (0.2, 0.1, 0.01)):
        v, t, p = tetra_mesh.make_tetrahedral_cylinder(r, L, hint)
        n = max(3, math.ceil(2.0 * np.pi * r / hint))
        V = tetra_mesh.tetrahedral_mesh_volumes(v[t])
        assert V.min() > 1e-12 * r * r * L
        assert np.sum(V) == pytest.approx(0.5 * n * r * r * np.sin(2 * np.pi / n) * L, rel=1e-12)
        np.testing.assert_array_almost_equal(tetra_mesh.center_of_mass_tetrahedral_mesh(v[t]), np.zeros(3))
        assert np.all(p[:2 + 2 * n] == 0.0) and np.all(p[2 + 2 * n:] > 0.0)
    for r, h, hint in ((0.1, 1.0, 0.01), (0.1, 0.0, 0.01), (0.2, 0.5, 0.02)):
        v, t, p = tetra_mesh.make_tetrahedral_capsule(r, h, hint)
        V = tetra_mesh.tetrahedral_mesh_volumes(v[t])
        # height 0: the medial segment is a point and the side prisms are
        # flat, as the reference builds them
        assert V.min() > 0.0 if h > 0.0 else V.min() >= 0.0
        assert np.sum(V) == pytest.approx(np.pi * r * r * h + 4.0 / 3.0 * np.pi * r ** 3, rel=2e-2)
        np.testing.assert_array_almost_equal(tetra_mesh.center_of_mass_tetrahedral_mesh(v[t]), np.zeros(3))
    rb = hc.RigidBody.make_cylinder(np.eye(4), 0.1, 1.0, resolution_hint=0.01)
    assert np.sum(tetra_mesh.tetrahedral_mesh_volumes(rb.tetrahedra_points)) == pytest.approx(
        np.pi * 0.01 * 1.0, abs=1e-2)


def test_missing_transform_is_a_key_error():
    """Transforms given as dicts are looked up per pair, as the reference does
    (_tetrahedron_intersection.py:65-66): a missing entry raises KeyError
    before any GPU work."""
    T = random_tets(np.random.RandomState(17), 3)
    e = np.ones((3, 4))
    X = {0: np.eye(4), 2: np.eye(4)}
    with pytest.raises(KeyError):
        hc.intersect_tetrahedron_pairs([(0, 0), (1, 2)], T, T, e, e, X, X)
    with pytest.raises(KeyError):
        hc.intersect_tetrahedron_pairs([(0, 0), (2, 1)], T, T, e, e, X, X)
    # the first missing key in pair order, X1[i] before X2[j]: pairs (0, 5),
    # (3, 0) with X1 missing 3 and X2 missing 5 raise for 5
    T6 = random_tets(np.random.RandomState(18), 6)
    e6 = np.ones((6, 4))
    X1 = {k: np.eye(4) for k in (0, 1, 2, 4, 5)}
    X2 = {k: np.eye(4) for k in (0, 1, 2, 3, 4)}
    with pytest.raises(KeyError) as info:
        hc.intersect_tetrahedron_pairs([(0, 5), (3, 0)], T6, T6, e6, e6, X1, X2)
    assert info.value.args == (5,)
    with pytest.raises(KeyError) as info:
        hc.intersect_tetrahedron_pairs([(3, 5)], T6, T6, e6, e6, X1, X2)
    assert info.value.args == (3,)


def test_rigid_body_transforms():  # test_hydroelastic_contact.py:67-86
    cube12origin = np.eye(4)
    cube1_22origin = np.copy(cube12origin)
    cube1_22origin[:3, 3] += 1.0
    rb1 = hc.RigidBody.make_cube(cube12origin, 0.1)
    cube22origin = np.eye(4)
    cube22origin[:3, 3] = [0.0, 0.0, 0.08]
    rb2 = hc.RigidBody.make_cube(cube22origin, 0.1)
    rb1.express_in(rb2.body2origin_)
    t1 = rb1.tetrahedra_points
    rb1.express_in(cube12origin)
    rb1.body2origin_ = cube1_22origin
    rb1.express_in(rb2.body2origin_)
    np.testing.assert_array_almost_equal(t1 + 1, rb1.tetrahedra_points)


# ------------------------------------------------------------------ GPU parity

@pytest.mark.gpu
def test_tet_mesh_prep_matches_oracle(gpu):
    from distance3d_amd import kernels
    rs = np.random.RandomState(5)
    T = random_tets(rs, 20000)
    T[7, 2] = T[7, 0]
    T[7, 3] = T[7, 1]  # degenerate
    ab, X, flags = kernels.tet_mesh_prep(T)
    want = oracle.tet_mesh_prep(T)
    assert np.array_equal(ab.cpu().numpy(), want[0])
    assert np.array_equal(flags.cpu().numpy(), want[2]) and want[2][7] == 1
    ok = want[2] == 0
    assert np.array_equal(X.cpu().numpy()[ok], want[1][ok])


@pytest.mark.gpu
def test_tet_contact_forces_match_oracle(gpu):
    from distance3d_amd import kernels
    rs = np.random.RandomState(6)
    n1, n2, m = 300, 400, 30000
    T1, T2 = random_tets(rs, n1), random_tets(rs, n2)
    e1, e2 = rs.uniform(0.0, 1.0, (n1, 4)), rs.uniform(0.0, 1.0, (n2, 4))
    e1[::5] = [0.0, 0.0, 0.0, 1.0]
    _, X1, _ = oracle.tet_mesh_prep(T1)
    _, X2, _ = oracle.tet_mesh_prep(T2)
    pairs = np.stack([rs.randint(0, n1, m), rs.randint(0, n2, m)], 1)
    pairs[:50, 1] = pairs[:50, 0]  # same-tetrahedron rows when T2 == T1 below
    for E1, E2 in ((1.0, 1.0), (4.0, 0.25)):
        st, pl, nv, po = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs, E1, E2)
        want = oracle.tet_contact_forces(T1, e1, pairs, st, pl, nv, po, E1)
        got = kernels.tet_contact_forces(T1, e1, pairs, st, pl, nv, po, E1).cpu().numpy()
        assert np.array_equal(got, want)
        assert 0.01 < np.mean(st & 1) < 0.9
    st, pl, nv, po = oracle.tet_pairs(T1, e1, X1, T1, e1, X1, pairs[:50])
    want = oracle.tet_contact_forces(T1, e1, pairs[:50], st, pl, nv, po)
    assert np.all(st & 0x10)
    assert np.array_equal(kernels.tet_contact_forces(T1, e1, pairs[:50], st, pl, nv, po).cpu().numpy(), want)


def _assert_surface_equal(got, want):
    assert got.intersection == want.intersection
    assert got.intersecting_tetrahedra1 == want.intersecting_tetrahedra1
    assert got.intersecting_tetrahedra2 == want.intersecting_tetrahedra2
    assert np.array_equal(got.contact_planes, want.contact_planes)
    assert len(got.contact_polygons) == len(want.contact_polygons)
    for a, b in zip(got.contact_polygons, want.contact_polygons):
        assert np.array_equal(a, b)
    assert np.array_equal(got.contact_areas, want.contact_areas)
    assert np.array_equal(got.contact_coms, want.contact_coms)
    assert np.array_equal(got.contact_forces, want.contact_forces)


@pytest.mark.gpu
def test_find_contact_surface_matches_oracle(gpu):
    """Spheres, ellipsoids, cubes and boxes in random overlapping poses: the
    GPU pipeline's contact surface equals the oracle pipeline's bit for bit
    (the tree broad phase gives the same pairs in the same order)."""
    rs = np.random.RandomState(7)
    makers = [lambda P: hc.RigidBody.make_sphere(P[:3, 3], 0.3, 2),
              lambda P: hc.RigidBody.make_ellipsoid(P, np.array([0.3, 0.2, 0.25]), 2),
              lambda P: hc.RigidBody.make_cube(P, 0.4),
              lambda P: hc.RigidBody.make_box(P, np.array([0.5, 0.3, 0.4]))]
    hits = 0
    for trial in range(12):
        a, b = makers[trial % 4], makers[(trial * 3 + 1) % 4]
        Pa, Pb = pose(rs, rs.uniform(-0.1, 0.1, 3)), pose(rs, rs.uniform(-0.3, 0.3, 3))
        Pa[:3, 3] += 0.05
        for trees in (False, True):
            got = hc.find_contact_surface(a(Pa), b(Pb), use_aabb_trees=trees)
            want = oracle_contact_surface(a(Pa), b(Pb))
            _assert_surface_equal(got, want)
        hits += got.intersection
    assert hits >= 8


@pytest.mark.gpu
def test_contact_forces_dropin_golden(gpu):
    """The reference's contact_forces cases (test_hydroelastic_contact.py:
    7-64, 256-293) through the GPU pipeline."""
    def pair(z):
        return (hc.RigidBody.make_sphere(np.array([0.0, 0.0, 0.01]), 0.15, 1),
                hc.RigidBody.make_sphere(np.array([0.0, 0.0, z]), 0.15, 1))

    for z, fz in ((0.4, None), (0.3, 1.120994e-06), (0.15, 0.00166866)):
        hit, w12, w21 = hc.contact_forces(*pair(z))
        want = oracle_contact_forces(*pair(z))
        assert hit == want[0]
        assert np.array_equal(w12, want[1]) and np.array_equal(w21, want[2])
        if fz is None:
            assert not hit
            np.testing.assert_array_almost_equal(w12, np.zeros(6))
            continue
        np.testing.assert_array_almost_equal(w12, [0.0, 0.0, fz, 0.0, 0.0, 0.0])
        np.testing.assert_array_almost_equal(w21, [0.0, 0.0, -fz, 0.0, 0.0, 0.0])
    hit, w12, w21, details = hc.contact_forces(*pair(0.3), return_details=True)
    n = len(details["contact_polygons"])
    assert hit and n > 0
    for key in ("contact_polygon_triangles", "contact_planes", "intersecting_tetrahedra1",
                "intersecting_tetrahedra2", "contact_coms", "contact_forces", "contact_areas",
                "pressures"):
        assert len(details[key]) == n
    assert "contact_point" in details
    rb1 = hc.RigidBody.make_sphere(np.array([0.5, 0, 0]), 0.15, 1)
    rb2 = hc.RigidBody.make_sphere(np.array([0.5, 0.12, 0]), 0.15, 1)
    i1, w12_1, w21_1, _ = hc.contact_forces(rb1, rb2, return_details=True)
    rb1.youngs_modulus, rb2.youngs_modulus = 1, 1
    i2, w12_2, w21_2, _ = hc.contact_forces(rb1, rb2, return_details=True)
    assert i1 == i2
    np.testing.assert_array_almost_equal(w12_1, w12_2)
    rb1.youngs_modulus, rb2.youngs_modulus = 10, 0.1
    i3, w12_3, _, _ = hc.contact_forces(rb1, rb2, return_details=True)
    assert i3
    np.testing.assert_array_almost_equal(w12_3, [0, 0.00032, 0, 0, 0, 0], decimal=4)
    assert np.linalg.norm(w12_3) < np.linalg.norm(w12_2)


@pytest.mark.gpu
def test_cylinder_tetrahedron_with_itself(gpu):  # test_hydroelastic_contact.py:89-98
    cyl = hc.RigidBody.make_cylinder(np.eye(4), 0.1, 1.0, resolution_hint=0.1)
    tet = cyl.tetrahedra_points[0]
    eps = np.array([0.0, 0.0, 0.0, 1.0])
    X = hc.barycentric_transforms(tet.reshape(1, 4, 3))[0]
    hit, _ = hc.intersect_tetrahedron_pair(tet, eps, X, tet, eps, X)
    assert hit


@pytest.mark.gpu
def test_single_polygon_and_mesh_helpers(gpu):
    """compute_contact_force on the same-tetrahedron case
    (test_hydroelastic_contact.py:101-116) and on random polygons against the
    oracle; tetrahedral_mesh_aabbs (:185-193) and barycentric_transforms."""
    sphere = hc.RigidBody.make_sphere(np.array([0.1, 0.2, 0.1]), 1.0, order=2)
    tet = sphere.tetrahedra_points[0]
    eps = np.array([0.0, 0.0, 0.0, 1.0])
    X = hc.barycentric_transforms(tet.reshape(1, 4, 3))[0]
    hit, (plane, poly) = hc.intersect_tetrahedron_pair(tet, eps, X, tet, eps, X)
    com, force, area, tris = hc.compute_contact_force(tet, eps, plane, poly)
    assert hit and area == 0.0
    np.testing.assert_array_almost_equal(force, np.zeros(3))
    rs = np.random.RandomState(11)
    T1, T2 = random_tets(rs, 200), random_tets(rs, 200)
    e1, e2 = rs.uniform(0, 1, (200, 4)), rs.uniform(0, 1, (200, 4))
    X1, X2 = hc.barycentric_transforms(T1), hc.barycentric_transforms(T2)
    assert np.array_equal(X1, oracle.tet_mesh_prep(T1)[1])
    pairs = np.stack([np.arange(200)] * 2, 1)
    st, pl, nv, po = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs, 3.0, 1.0)
    F = oracle.tet_contact_forces(T1, e1, pairs, st, pl, nv, po, 3.0)
    for k in np.nonzero(st & 1)[0][:10]:
        com, force, area, tris = hc.compute_contact_force(T1[k], e1[k], pl[k], po[k, :nv[k]], 3.0)
        assert np.array_equal(np.concatenate([com, force, [area]]), F[k])
        assert np.array_equal(tris, hc.TRIANGLES[:nv[k] - 2])
    rb = hc.RigidBody.make_sphere(np.array([0.0, 0.2, -0.3]), 0.2, 2)
    rb.express_in(np.eye(4))
    ab = hc.tetrahedral_mesh_aabbs(rb.tetrahedra_points)
    assert np.array_equal(ab[:, :, 0], rb.tetrahedra_points.min(axis=1))
    assert np.array_equal(ab[:, :, 1], rb.tetrahedra_points.max(axis=1))


class _Frames:
    """Poses of frames in "origin" (the transform-manager calls the BVH makes)."""

    def __init__(self):
        self.T = {}

    def add_transform(self, a, b, A2B):
        self.T[a] = np.asarray(A2B, dtype=float)

    def get_transform(self, a, b):
        return self.T[a]


@pytest.mark.gpu
def test_hydroelastic_bvh(gpu):
    """_broad_phase.py:7-98: body boxes from the tetrahedron boxes; overlap
    queries equal brute force over those boxes."""
    rs = np.random.RandomState(15)
    tm = _Frames()
    bvh = hc.HydroelasticBoundingVolumeHierarchy(tm, "base")
    bodies = {}
    for k in range(12):
        P = pose(rs, rs.uniform(-1.0, 1.0, 3))
        tm.add_transform("b%d" % k, "origin", P)
        body = (hc.RigidBody.make_sphere(P[:3, 3], rs.uniform(0.1, 0.5), 1) if k % 2 else
                hc.RigidBody.make_box(P, rs.uniform(0.2, 0.8, 3)))
        body.express_in(np.eye(4))
        bodies["b%d" % k] = body
        bvh.add_collider("b%d" % k, body)
    boxes = {f: b.aabb() for f, b in bodies.items()}
    for f, b in bodies.items():
        pts = b.tetrahedra_points.reshape(-1, 3)
        assert np.array_equal(boxes[f][:, 0], pts.min(axis=0))
        assert np.array_equal(boxes[f][:, 1], pts.max(axis=0))

    def overlap(a, b):
        return bool(np.all(a[:, 0] <= b[:, 1]) and np.all(a[:, 1] >= b[:, 0]))

    for f, b in bodies.items():
        got = set(bvh.aabb_overlapping_colliders(b, whitelist=(f,)))
        want = {g for g in bodies if g != f and overlap(boxes[g], boxes[f])}
        assert got == want
    pairs = {tuple(sorted((p[0][0], p[1][0]))) for p in bvh.aabb_overlapping_with_self()}
    want = {tuple(sorted((f, g))) for f in bodies for g in bodies if f != g and overlap(boxes[f], boxes[g])}
    assert pairs == want and len(want) > 0


@pytest.mark.gpu
def test_empty_and_disjoint_inputs(gpu):
    """Bodies far apart (no candidate pairs), candidate pairs with no hit,
    and empty arrays through every new wrapper."""
    from distance3d_amd import kernels
    rb1 = hc.RigidBody.make_sphere(np.array([0.0, 0.0, 0.0]), 0.15, 1)
    rb2 = hc.RigidBody.make_sphere(np.array([5.0, 0.0, 0.0]), 0.15, 1)
    s = hc.find_contact_surface(rb1, rb2)
    assert not s.intersection and len(s.contact_polygons) == 0 and s.contact_planes.shape == (0, 4)
    hit, w12, w21 = hc.contact_forces(hc.RigidBody.make_sphere(np.zeros(3), 0.15, 1),
                                      hc.RigidBody.make_sphere(np.array([5.0, 0, 0]), 0.15, 1))
    assert not hit and np.all(w12 == 0.0) and np.all(w21 == 0.0)
    assert hc.contact_forces(hc.RigidBody.make_sphere(np.zeros(3), 0.15, 1),
                             hc.RigidBody.make_sphere(np.array([5.0, 0, 0]), 0.15, 1),
                             return_details=True)[3] == {}
    # pairs of tetrahedra far apart: no contact polygon
    rs = np.random.RandomState(16)
    T = random_tets(rs, 50)
    z = rs.uniform(0.1, 1.0, (50, 4))
    X = hc.barycentric_transforms(T)
    res = hc.intersect_tetrahedron_pairs([(i, i) for i in range(50)], T, T + 10.0, z, z, X,
                                         hc.barycentric_transforms(T + 10.0))
    assert res[0] is False and res[3] == [] and res[4] == []
    assert hc.intersect_tetrahedron_pairs([], T, T, z, z, X, X) == (False, [], [], [], [])
    ab, Xe, fl = kernels.tet_mesh_prep(np.empty((0, 4, 3)))
    assert ab.shape == (0, 3, 2) and Xe.shape == (0, 4, 4) and fl.shape == (0,)
    assert kernels.polygon_angles(np.empty(0), np.empty(0)).shape == (0,)
    out = kernels.tet_contact_forces(T, z, np.empty((0, 2)), np.empty(0), np.empty((0, 4)),
                                     np.empty(0), np.empty((0, kernels.TET_MAX_POLY, 3)))
    assert out.shape == (0, 7)
