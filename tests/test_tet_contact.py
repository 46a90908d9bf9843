"""Hydroelastic tetrahedron-pair contact (hydroelastic_contact/
_tetrahedron_intersection.py, _halfplanes.py): the oracle pinned by the
reference's own cases and by what the geometry fixes (the polygon lies on the
plane, inside both tetrahedra, where the weighted potentials agree, and has
the area of the plane section of one tetrahedron clipped by the other); on
the GPU the batched kernel against the oracle."""
import numpy as np
import pytest

from oracle import oracle


def barycentric_transforms(T):
    """_barycentric_transform.py:4-9: X with X (r, 1) = barycentric coords."""
    T = np.asarray(T, dtype=np.float64).reshape(-1, 4, 3)
    return np.linalg.pinv(np.concatenate([T.transpose(0, 2, 1), np.ones((len(T), 1, 4))], axis=1))


def random_tets(rs, n, spread=0.6):
    c = rs.uniform(-spread, spread, (n, 1, 3))
    return c + rs.randn(n, 4, 3) * 0.5


def _section(T, n, d):
    """Polygon of the plane n.x = d inside tetrahedron T (edge crossings,
    ordered around their centroid)."""
    pts = []
    s = T @ n - d
    for a in range(4):
        for b in range(a + 1, 4):
            if (s[a] < 0) != (s[b] < 0):
                t = s[a] / (s[a] - s[b])
                pts.append(T[a] + t * (T[b] - T[a]))
    return _order(np.array(pts), n)


def _order(P, n):
    if len(P) < 3:
        return P
    u = np.cross(n, [1.0, 0.0, 0.0])
    if np.linalg.norm(u) < 1e-6:
        u = np.cross(n, [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    c = P.mean(0)
    return P[np.argsort(np.arctan2((P - c) @ v, (P - c) @ u))]


def _clip(P, T):
    """Sutherland-Hodgman: convex polygon P clipped by the 4 half-spaces of
    tetrahedron T (inside = the side of the opposite vertex)."""
    for f in range(4):
        idx = [k for k in range(4) if k != f]
        a, b, c = T[idx]
        nrm = np.cross(b - a, c - a)
        if np.dot(T[f] - a, nrm) < 0:
            nrm = -nrm
        out = []
        for k in range(len(P)):
            p, q = P[k], P[(k + 1) % len(P)]
            sp, sq = np.dot(p - a, nrm), np.dot(q - a, nrm)
            if sp >= 0:
                out.append(p)
            if (sp >= 0) != (sq >= 0):
                out.append(p + sp / (sp - sq) * (q - p))
        P = np.array(out)
        if len(P) == 0:
            break
    return P


def _area(P, n):
    if len(P) < 3:
        return 0.0
    return 0.5 * abs(sum(np.dot(n, np.cross(P[k], P[(k + 1) % len(P)])) for k in range(len(P))))


def test_oracle_intersect_halfplanes_known_answer():  # test_hydroelastic_contact.py:119-136
    halfplanes = np.array([[1.0, 0.0, 0.0, 1.0], [0.5, 0.0, 0.0, 1.0], [0.0, 0.5, -1.0, 0.0],
                           [-0.5, 0.0, 0.0, -1.0], [0.0, -0.5, 1.0, 0.0]])
    polygon = [p.tolist() for p in oracle.intersect_halfplanes(halfplanes)]
    for p in ([0.5, 0.5], [-0.5, 0.5], [-0.5, -0.5], [0.5, -0.5]):
        assert p in polygon


def test_oracle_same_tetrahedron():  # test_hydroelastic_contact.py:89-98
    rs = np.random.RandomState(3)
    T = random_tets(rs, 1)
    X = barycentric_transforms(T)
    e = np.array([[0.0, 0.0, 0.0, 1.0]])
    st, plane, nv, poly = oracle.tet_pairs(T, e, X, T, e, X, [[0, 0]])
    assert st[0] & 1 and st[0] & 0x10
    # the plane passes through the potential-weighted centre (= vertex 3)
    assert np.allclose(poly[0, 0], T[0, 3]) and np.allclose(plane[0, :3] * plane[0, 3], T[0, 3])


def test_oracle_polygon_geometry():
    """On random overlapping pairs: polygon on the contact plane, inside both
    tetrahedra, weighted potentials equal on it, area equal to the clipped
    plane section; and no polygon where the brute-force section is empty."""
    rs = np.random.RandomState(4)
    n = 400
    T1, T2 = random_tets(rs, n), random_tets(rs, n)
    e1, e2 = rs.uniform(0.0, 1.0, (n, 4)), rs.uniform(0.0, 1.0, (n, 4))
    X1, X2 = barycentric_transforms(T1), barycentric_transforms(T2)
    E1, E2 = 1.0, 2.5
    pairs = np.stack([np.arange(n), np.arange(n)], 1)
    st, planes, nv, polys = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs, E1, E2)
    hits = 0
    for k in range(n):
        nrm, d = planes[k, :3], planes[k, 3]
        assert abs(np.linalg.norm(nrm) - 1.0) < 1e-12
        bf = _area(_clip(_section(T1[k], nrm, d), T2[k]), nrm)
        if not st[k] & 1:
            assert bf < 1e-6
            continue
        hits += 1
        P = polys[k, :nv[k]]
        assert np.all(np.abs(P @ nrm - d) < 1e-9)
        h = np.concatenate([P, np.ones((len(P), 1))], 1)
        b1, b2 = h @ X1[k].T, h @ X2[k].T
        assert np.all(b1 > -1e-9) and np.all(b2 > -1e-9)
        assert np.allclose(E1 * (b1 @ e1[k]), E2 * (b2 @ e2[k]), rtol=0, atol=1e-9)
        assert _area(P, nrm) == pytest.approx(bf, rel=1e-9, abs=1e-12)
    assert 20 < hits < n - 20


def test_oracle_plane_sign_matters():
    """The checks above discriminate: the plane with d negated puts the
    polygon off the plane."""
    rs = np.random.RandomState(5)
    n = 100
    T1, T2 = random_tets(rs, n), random_tets(rs, n)
    e1, e2 = rs.uniform(0.0, 1.0, (n, 4)), rs.uniform(0.0, 1.0, (n, 4))
    pairs = np.stack([np.arange(n), np.arange(n)], 1)
    st, planes, nv, polys = oracle.tet_pairs(T1, e1, barycentric_transforms(T1), T2, e2,
                                             barycentric_transforms(T2), pairs)
    k = int(np.nonzero((st & 1) & (np.abs(planes[:, 3]) > 1e-3))[0][0])
    P = polys[k, :nv[k]]
    assert np.any(np.abs(P @ planes[k, :3] + planes[k, 3]) > 1e-6)


def _angle_inputs(rs, n):
    y = rs.randn(n) * np.exp(rs.uniform(-20, 20, n))
    x = rs.randn(n) * np.exp(rs.uniform(-20, 20, n))
    h = n // 4
    y[:h] = x[:h] * (1 + rs.randn(h) * 1e-12)  # near the diagonals
    y[h:2 * h] = rs.uniform(-1, 1, h)
    x[h:2 * h] = rs.uniform(-1, 1, h)
    sp = [0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 1e-310, -1e-310, 1e300]
    yy, xx = np.meshgrid(sp, sp)
    return np.concatenate([y, yy.ravel()]), np.concatenate([x, xx.ravel()])


def test_oracle_angle_is_correctly_rounded():
    """The ordering's atan2 against 50-digit mpmath on a sample (including
    near-diagonal third-quadrant inputs, where np.arctan2 and the C library
    misround on some hosts) and against np.arctan2's special cases."""
    mpmath = pytest.importorskip("mpmath")
    rs = np.random.RandomState(12)
    y, x = _angle_inputs(rs, 4000)
    got = oracle.atan2(y, x)
    sp = slice(4000, None)
    assert np.array_equal(got[sp], np.arctan2(y[sp], x[sp]), equal_nan=True)
    assert np.array_equal(np.signbit(got[sp]), np.signbit(np.arctan2(y[sp], x[sp])))
    mpmath.mp.dps = 50
    for k in range(4000):
        assert got[k] == float(mpmath.atan2(mpmath.mpf(y[k]), mpmath.mpf(x[k]))), (y[k], x[k])


@pytest.mark.gpu
def test_tet_pairs_match_oracle(gpu):
    """Random pairs over two bodies (overlapping and not, shared
    tetrahedra, both moduli): status, plane, vertex count and polygon equal
    the oracle's bit for bit."""
    from distance3d_amd import kernels
    rs = np.random.RandomState(8)
    n1, n2, m = 300, 400, 20000
    T1, T2 = random_tets(rs, n1), random_tets(rs, n2)
    e1, e2 = rs.uniform(0.0, 1.0, (n1, 4)), rs.uniform(0.0, 1.0, (n2, 4))
    e1[::7] = [0.0, 0.0, 0.0, 1.0]
    X1, X2 = barycentric_transforms(T1), barycentric_transforms(T2)
    pairs = np.stack([rs.randint(0, n1, m), rs.randint(0, n2, m)], 1)
    for E1, E2 in ((1.0, 1.0), (3.0, 0.5)):
        got = [t.cpu().numpy() for t in kernels.tet_pairs_batch(T1, e1, X1, T2, e2, X2, pairs, E1, E2)]
        want = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs, E1, E2)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert np.array_equal(got[2], want[2])
        assert np.array_equal(got[3], want[3], equal_nan=True)
        assert 0.01 < np.mean(want[0] & 1) < 0.9
    # same tetrahedron on both sides
    got = kernels.tet_pairs_batch(T1, e1, X1, T1, e1, X1, np.stack([np.arange(n1)] * 2, 1))
    want = oracle.tet_pairs(T1, e1, X1, T1, e1, X1, np.stack([np.arange(n1)] * 2, 1))
    assert np.all(want[0] & 0x10)
    assert np.array_equal(got[0].cpu().numpy(), want[0])
    assert np.array_equal(got[3].cpu().numpy(), want[3], equal_nan=True)


@pytest.mark.gpu
def test_tet_dropin(gpu):
    """intersect_tetrahedron_pair / _pairs (the reference's surface):
    test_hydroelastic_contact.py:89-98's same-tetrahedron case, and the
    batched call returns the intersecting pairs in order with the oracle's
    planes and polygons."""
    from distance3d_amd import hydroelastic_contact as hc
    rs = np.random.RandomState(9)
    T = random_tets(rs, 1)[0]
    e = np.array([0.0, 0.0, 0.0, 1.0])
    X = hc.barycentric_transforms(T.reshape(1, 4, 3))[0]
    hit, (plane, poly) = hc.intersect_tetrahedron_pair(T, e, X, T, e, X)
    assert hit and poly.shape == (3, 3)
    n = 200
    T1, T2 = random_tets(rs, n), random_tets(rs, n)
    e1, e2 = rs.uniform(0.0, 1.0, (n, 4)), rs.uniform(0.0, 1.0, (n, 4))
    X1, X2 = hc.barycentric_transforms(T1), hc.barycentric_transforms(T2)
    pairs = [(i, (7 * i) % n) for i in range(n)]
    X1d = {i: X1[i] for i in range(n)}
    X2d = {j: X2[j] for j in range(n)}
    inter, planes, polys, it1, it2 = hc.intersect_tetrahedron_pairs(pairs, T1, T2, e1, e2, X1d, X2d)
    st, wp, wnv, wpoly = oracle.tet_pairs(T1, e1, X1, T2, e2, X2, pairs)
    hits = np.nonzero(st & 1)[0]
    assert inter and it1 == [pairs[k][0] for k in hits] and it2 == [pairs[k][1] for k in hits]
    assert np.array_equal(planes, wp[hits])
    for p, k in zip(polys, hits):
        assert np.array_equal(p, wpoly[k, :wnv[k]])
    miss = int(np.nonzero((st & 1) == 0)[0][0])
    i, j = pairs[miss]
    hit, (plane, poly) = hc.intersect_tetrahedron_pair(T1[i], e1[i], X1[i], T2[j], e2[j], X2[j])
    assert not hit and np.array_equal(plane, wp[miss])
    assert (poly is None) == (not st[miss] & 0x20)


@pytest.mark.gpu
def test_polygon_angles_match_oracle(gpu):
    """The kernels' atan2 equals the oracle's bit for bit."""
    from distance3d_amd import kernels
    rs = np.random.RandomState(13)
    y, x = _angle_inputs(rs, 1 << 20)
    got = kernels.polygon_angles(y, x).cpu().numpy()
    want = oracle.atan2(y, x)
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(np.signbit(got), np.signbit(want))


@pytest.mark.gpu
def test_intersect_halfplanes_dropin(gpu):
    """_halfplanes.py:29-64 on the GPU: the reference's known answer
    (test_hydroelastic_contact.py:119-136), random sets against the oracle
    bit for bit (order included), and the overflow where the reference
    asserts (eight lines through one point: 28 valid intersections)."""
    from distance3d_amd import hydroelastic_contact as hc
    halfplanes = np.array([[1.0, 0.0, 0.0, 1.0], [0.5, 0.0, 0.0, 1.0], [0.0, 0.5, -1.0, 0.0],
                           [-0.5, 0.0, 0.0, -1.0], [0.0, -0.5, 1.0, 0.0]])
    got = hc.intersect_halfplanes(halfplanes)
    assert np.array_equal(got, oracle.intersect_halfplanes(halfplanes))
    for p in ([0.5, 0.5], [-0.5, 0.5], [-0.5, -0.5], [0.5, -0.5]):
        assert p in got.tolist()
    rs = np.random.RandomState(18)
    for n in (0, 1, 2, 3, 5, 8, 13, 70):
        ang = rs.uniform(0, 2 * np.pi, n)
        pts = rs.uniform(-0.3, 0.3, (n, 2)) - 0.5 * np.stack([np.cos(ang + np.pi / 2), np.sin(ang + np.pi / 2)], 1)
        hp = np.concatenate([pts, np.stack([np.cos(ang), np.sin(ang)], 1)], 1)
        try:
            want = oracle.intersect_halfplanes(hp)
        except AssertionError:
            with pytest.raises(AssertionError):
                hc.intersect_halfplanes(hp)
            continue
        assert np.array_equal(hc.intersect_halfplanes(hp), want)
    ang = np.arange(8) * np.pi / 8
    star = np.stack([np.zeros(8), np.zeros(8), np.cos(ang), np.sin(ang)], 1)
    with pytest.raises(AssertionError):
        oracle.intersect_halfplanes(star)
    with pytest.raises(AssertionError):
        hc.intersect_halfplanes(star)


def test_polygon_centre_summation_order_does_not_change_the_order():
    """order_points (_tetrahedron_intersection.py:299-318) sorts a contact
    polygon's points by angle around their mean.  The oracle and the kernel
    sum the mean one element after another, as numba's np.mean does; numpy's
    own np.mean sums pairwise from 8 elements on, and the two centres can
    differ in the last bits.  For convex polygons of 8 to 23 points in general
    position the angular order around either centre is the same, so the
    summation order cannot change the ordered polygon there."""
    rs = np.random.RandomState(17)
    differ = 0
    for _ in range(3000):
        n = rs.randint(8, 24)
        t = np.sort(rs.uniform(-np.pi, np.pi, n))
        axes = rs.uniform(0.1, 2.0, 2)
        pts = np.column_stack([axes[0] * np.cos(t), axes[1] * np.sin(t)])
        a = rs.uniform(0, 2 * np.pi)
        R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        pts = pts @ R.T + rs.uniform(-5, 5, 2)
        pts = pts[rs.permutation(n)]
        seq = np.zeros(2)
        for p in pts:
            seq = seq + p
        seq = seq / n
        pair = np.array([np.mean(pts[:, 0]), np.mean(pts[:, 1])])
        differ += int(np.any(seq != pair))
        a_seq = np.arctan2(pts[:, 1] - seq[1], pts[:, 0] - seq[0])
        a_pair = np.arctan2(pts[:, 1] - pair[1], pts[:, 0] - pair[0])
        assert np.array_equal(np.argsort(a_seq, kind="stable"), np.argsort(a_pair, kind="stable"))
    assert differ > 300  # the centres do differ, often
