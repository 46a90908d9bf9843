"""EPA with the reference's call surface (epa.py:9-78), evaluated by the
gfx950 EPA kernel (one wavefront per pair, polytope in LDS).

``epa(simplex, collider1, collider2)`` takes the (4, 3) simplex returned by
``gjk.gjk`` (or any (4, 3) array).  Its rows are used as the seed
tetrahedron; when they do not span a tetrahedron (GJK stopped with fewer than
four support points) the seed is blown up with extra support points, so the
call is well defined for every intersecting pair.
"""
import numpy as np

from . import kernels
from .scene import SceneBuilder

epa_batch = kernels.epa_batch


def _spans_tetrahedron(Y):
    """True when the four rows are affinely independent (a non-flat
    tetrahedron): the reference's Polytope seeds its four faces from exactly
    these rows (epa.py:89-97)."""
    e = Y[1:] - Y[0]
    vol = abs(np.linalg.det(e))
    scale = np.prod(np.linalg.norm(e, axis=1))
    return bool(np.isfinite(vol) and vol > 1e-12 * scale and scale > 0.0)


def epa(simplex, collider1, collider2, max_iter=64, max_loose_edges=32, max_faces=64,
        epsilon=1e-8):
    """Minimum translation vector of two intersecting colliders (epa.py:9-78).

    Returns ``(mtv, faces, success)`` like the reference: ``mtv`` (3,) is the
    penetration depth times the contact normal (subtract it from collider 1
    or add it to collider 2 to separate them), ``faces`` (n_faces, 4, 3) is
    the final polytope (three vertices and the outward unit normal per face,
    the reference's ``Polytope.get_all_faces``), ``success`` is True when the
    expansion converged within ``max_iter`` iterations.

    Seeding:
      - the simplex ``gjk.gjk`` returned carries the support points on both
        colliders and its number of valid rows; EPA starts from those rows
        (blown up to a tetrahedron when GJK ended with fewer than four);
      - a plain (4, 3) array whose rows span a tetrahedron is used as given,
        as the reference does;
      - a plain array whose rows are flat cannot seed a polytope (the
        reference's face normals would be zero); GJK is run on the GPU for
        the pair to obtain a seed.

    The polytope holds at most 128 faces (``max_faces`` above that is
    clamped).  Running out of faces raises ``AssertionError``, as the
    reference's ``assert n_faces < max_faces`` does (epa.py:128).
    ``max_loose_edges`` is accepted for signature compatibility: horizon
    edges are matched by vertex index and need no cap.
    """
    b = SceneBuilder()
    i1 = collider1.pack(b)
    i2 = collider2.pack(b)
    scene = b.build().to_device()
    Y = np.asarray(simplex, dtype=np.float64).reshape(4, 3)
    support1 = getattr(simplex, "support1", None)
    if support1 is not None and getattr(simplex, "n_points", None) is not None:
        s = np.stack([Y, support1, simplex.support2])[None]
        n_points = [simplex.n_points]
    elif _spans_tetrahedron(Y):
        # the caller's rows; their split into points of collider 1 and 2 is
        # unknown, which only the contact points (not returned here) need
        s = np.stack([Y, Y, np.zeros((4, 3))])[None]
        n_points = [4]
    else:
        g = kernels.gjk_distance_batch(scene, [i1], [i2], want_simplex=True)
        s, n_points = g.simplex, g.n_points
    if int(max_iter) < 1:
        raise ValueError("max_iter must be >= 1")
    mf = min(int(max_faces), 128)
    r = kernels.epa_batch(scene, [i1], [i2], s, n_points, max_iterations=int(max_iter),
                          max_faces=mf, epsilon=epsilon, want_faces=True)
    mtv = r.mtv.cpu().numpy()[0]
    status = int(r.status.cpu().item())
    if status == 0x200:
        raise ValueError("invalid collider")
    if status == 2:
        raise AssertionError("EPA polytope reached max_faces=%d" % mf)
    nf = int(r.n_faces.cpu().item())
    faces = r.faces.cpu().numpy()[0, :nf].copy()
    return mtv, faces, status == 1


def epa_contact(collider1, collider2, max_iter=64, max_faces=128, epsilon=1e-8):
    """GJK then EPA for one pair: ``(depth, normal, contact_point1,
    contact_point2, success)``; depth 0 and success False when the colliders
    do not intersect."""
    b = SceneBuilder()
    i1 = collider1.pack(b)
    i2 = collider2.pack(b)
    scene = b.build().to_device()
    g = kernels.gjk_distance_batch(scene, [i1], [i2], want_simplex=True)
    if float(g.dist.cpu().item()) > 0.0:
        return 0.0, np.zeros(3), None, None, False
    r = kernels.epa_batch(scene, [i1], [i2], g.simplex, g.n_points, max_iterations=max_iter,
                          max_faces=max_faces, epsilon=epsilon)
    out = r.raw.cpu().numpy()[0]
    return out[0], out[1:4], out[7:10], out[10:13], int(r.status.cpu().item()) == 1
