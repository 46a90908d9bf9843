"""EPA with the reference's call surface (epa.py:9-78), evaluated by the
gfx950 EPA kernel (one wavefront per pair, polytope in LDS).

``epa(simplex, collider1, collider2)`` takes the (4, 3) simplex returned by
``gjk.gjk``.  Its rows are used as the seed tetrahedron; when they do not
span a tetrahedron (GJK stopped with fewer than four support points) the seed
is blown up with extra support points, so the call is well defined for every
intersecting pair.
"""
import numpy as np

from . import kernels
from .scene import SceneBuilder

epa_batch = kernels.epa_batch


def epa(simplex, collider1, collider2, max_iter=64, max_loose_edges=32, max_faces=64,
        epsilon=1e-8):
    """Minimum translation vector of two intersecting colliders.

    Returns ``(mtv, faces, success)`` like the reference: ``mtv`` (3,) is the
    penetration depth times the contact normal (subtract it from collider 1
    or add it to collider 2 to separate them), ``faces`` (n_faces, 4, 3) is
    not materialised by the GPU kernel and is returned empty, ``success`` is
    True when the expansion converged.

    ``simplex`` is the one ``gjk.gjk`` returned for these colliders; when it
    is a plain array without the support points GJK recorded, GJK is rerun
    on the GPU to obtain them (the seed only changes the path, not the
    converged answer).  ``max_loose_edges`` is accepted for signature
    compatibility: horizon edges are matched by vertex index and need no
    cap.  ``max_faces`` may be at most 128.
    """
    b = SceneBuilder()
    i1 = collider1.pack(b)
    i2 = collider2.pack(b)
    scene = b.build().to_device()
    if getattr(simplex, "n_points", None) is not None and simplex.support1 is not None:
        s = np.stack([np.asarray(simplex, dtype=np.float64), simplex.support1,
                      simplex.support2])[None]
        n_points = [simplex.n_points]
    else:
        g = kernels.gjk_distance_batch(scene, [i1], [i2], want_simplex=True)
        s, n_points = g.simplex, g.n_points
    r = kernels.epa_batch(scene, [i1], [i2], s, n_points, max_iterations=max_iter,
                          max_faces=min(int(max_faces), 128), epsilon=epsilon)
    mtv = r.mtv.cpu().numpy()[0]
    status = int(r.status.cpu().item())
    return mtv, np.zeros((0, 4, 3)), status == 1


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
