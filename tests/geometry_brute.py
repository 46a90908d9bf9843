"""Independent brute-force / closed-form geometry used to pin the oracle.

Nothing here shares code with the oracle or the kernels: closest points come
from enumerating simplex faces (exact active-set search), distances between
hulls from scipy's Qhull on the Minkowski difference, segment distances from
the clamped closed form.
"""
import itertools

import numpy as np
from scipy.spatial import ConvexHull


def closest_point_of_points(points):
    """Closest point to the origin of conv(points), |points| <= 5, by trying
    every affinely independent subset and keeping feasible projections."""
    P = np.asarray(points, dtype=float)
    best = None
    for k in range(1, len(P) + 1):
        for sub in itertools.combinations(range(len(P)), k):
            S = P[list(sub)]
            if k == 1:
                lam = np.array([1.0])
            else:
                # minimise |S^T lam| s.t. sum lam = 1 (KKT system)
                G = S @ S.T
                A = np.zeros((k + 1, k + 1))
                A[:k, :k] = 2 * G
                A[:k, k] = 1
                A[k, :k] = 1
                rhs = np.zeros(k + 1)
                rhs[k] = 1
                try:
                    sol = np.linalg.solve(A, rhs)
                except np.linalg.LinAlgError:
                    continue
                lam = sol[:k]
                if np.any(lam < -1e-12):
                    continue
            x = lam @ S
            d = np.linalg.norm(x)
            if best is None or d < best[0] - 1e-15:
                best = (d, x)
    return best[1]


def point_triangle_distance(p, a, b, c):
    return np.linalg.norm(closest_point_of_points(np.array([a, b, c]) - p))


def hull_distance(v1, v2):
    """Distance between conv(v1) and conv(v2) and whether they intersect
    (signed: negative inside), via Qhull on the Minkowski difference."""
    md = (np.asarray(v1)[:, None, :] - np.asarray(v2)[None, :, :]).reshape(-1, 3)
    h = ConvexHull(md)
    offs = h.equations[:, :3] @ np.zeros(3) + h.equations[:, 3]
    if np.all(offs <= 0):
        return 0.0, float(np.max(offs))  # inside; max(offs) = -depth
    best = np.inf
    for s in h.simplices:
        best = min(best, point_triangle_distance(np.zeros(3), *md[s]))
    return best, float(np.max(offs))


def penetration_depth(v1, v2):
    """min over facets of hull(A - B) of the facet's distance to the origin
    (origin inside): depth and outward unit normal."""
    md = (np.asarray(v1)[:, None, :] - np.asarray(v2)[None, :, :]).reshape(-1, 3)
    h = ConvexHull(md)
    d = -h.equations[:, 3]
    k = int(np.argmin(d))
    return float(d[k]), h.equations[k, :3]


def segment_segment(p1, q1, p2, q2):
    """Closest points of segments p1q1 and p2q2 (clamped closed form)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c = d1 @ r
    b = d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-15 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    c1, c2 = p1 + d1 * s, p2 + d2 * t
    return np.linalg.norm(c1 - c2), c1, c2


def random_rotation(rs):
    q = rs.randn(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def random_pose(rs, scale=1.0):
    T = np.eye(4)
    T[:3, :3] = random_rotation(rs)
    T[:3, 3] = rs.randn(3) * scale
    return T
