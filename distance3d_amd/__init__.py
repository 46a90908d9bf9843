"""distance3d_amd: distance3d's collision hot path (GJK, EPA, collider
support functions, AABB broad phase) on AMD Instinct MI355X (gfx950).

Drop-in modules: ``gjk``, ``epa``, ``colliders``, ``containment``,
``aabb_tree``, ``broad_phase``.  Batched device-resident work:
``scene`` + ``kernels``.
"""
__version__ = "0.1.0"
