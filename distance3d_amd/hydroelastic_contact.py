"""Hydroelastic contact of tetrahedron pairs with the reference's call
surface (hydroelastic_contact/_tetrahedron_intersection.py:7-139,
_barycentric_transform.py), evaluated by the gfx950 kernel
(d3d_tet_pairs_batch): for each pair the contact plane where the two
Young's-modulus-weighted pressure fields agree, and the contact polygon.
"""
import numpy as np

from . import kernels

tet_pairs_batch = kernels.tet_pairs_batch

SAME = 0x10
CUTS = 0x20
OVERFLOW = 0x100


def barycentric_transforms(tetrahedra_points):
    """X per tetrahedron with X.dot((r, 1)) = barycentric coordinates of r
    (_barycentric_transform.py:4-9, the same pseudo-inverse)."""
    T = np.asarray(tetrahedra_points, dtype=np.float64)
    return np.linalg.pinv(np.hstack((T.transpose((0, 2, 1)), np.ones((len(T), 1, 4)))))


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
    st, planes, nv, polys = kernels.tet_pairs_batch(
        T1, epsilon1, _dense(X1, len(T1)), T2, epsilon2, _dense(X2, len(T2)), pr,
        youngs_modulus1, youngs_modulus2)
    st, planes, nv, polys = st.cpu().numpy(), planes.cpu().numpy(), nv.cpu().numpy(), polys.cpu().numpy()
    _check(st)
    hit = np.nonzero(st & 1)[0]
    contact_planes = planes[hit] if len(hit) else []
    return (len(hit) > 0, contact_planes, [polys[k, :nv[k]] for k in hit],
            pr[hit, 0].tolist(), pr[hit, 1].tolist())


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
