"""Bounding volume hierarchy of colliders (broad_phase.py:13-262) on top of
the GPU AABB tree.

``tm`` is any transform manager with ``add_transform(from, to, A2B)`` and
``get_transform(from, to)`` (pytransform3d's TransformManager has both);
filling the tree from a URDF needs pytransform3d and distance3d's mesh
loader and is only available when those are importable.
"""
import warnings

import numpy as np

from .aabb_tree import AabbTree


class BoundingVolumeHierarchy:
    """Bounding volume hierarchy (BVH) for broad phase collision detection."""

    def __init__(self, tm, base_frame, base_frame2origin=None):
        if base_frame2origin is None:
            base_frame2origin = np.eye(4)
        tm.add_transform(base_frame, "origin", base_frame2origin)
        self.tm = tm
        self.base_frame = base_frame
        self.base_frame2origin = base_frame2origin
        self.collider_frames = set()
        self.aabbtree_ = AabbTree()
        self.colliders_ = {}
        self.self_collision_whitelists_ = {}

    def fill_tree_with_colliders(self, tm, make_artists=False,
                                 fill_self_collision_whitelists=False, use_visuals=False):
        """Fill the tree from a URDF transform manager (needs pytransform3d)."""
        from pytransform3d import urdf  # noqa: F401  (raises ImportError when absent)
        from .colliders import Box, Cylinder, Sphere
        objects = tm.visuals if use_visuals else tm.collision_objects
        for obj in objects:
            try:
                A2B = tm.get_transform(obj.frame, "origin")
                if isinstance(obj, urdf.Sphere):
                    collider = Sphere(center=A2B[:3, 3], radius=obj.radius)
                elif isinstance(obj, urdf.Box):
                    collider = Box(A2B, obj.size)
                elif isinstance(obj, urdf.Cylinder):
                    collider = Cylinder(cylinder2origin=A2B, radius=obj.radius,
                                        length=obj.length)
                else:
                    raise RuntimeError("mesh colliders from URDF need distance3d.io")
                if make_artists:
                    collider.make_artist()
                self.add_collider(obj.frame, collider)
            except RuntimeError as e:
                warnings.warn(str(e))
        if fill_self_collision_whitelists:
            raise ImportError("self-collision whitelists need distance3d.urdf_utils")
        self.update_collider_poses()

    def add_collider(self, frame, collider):
        self.collider_frames.add(frame)
        self.colliders_[frame] = collider
        self.aabbtree_.insert_aabb(collider.aabb(), (frame, collider))

    def update_collider_poses(self):
        """Update poses of all colliders from the transform manager
        (broad_phase.py:144-151).  The reference rebuilds the tree; when the
        tree already holds exactly these colliders in this order, its leaves
        get the new boxes and the device tree is refitted instead."""
        frames = list(self.colliders_)
        aabbs = []
        for frame in frames:
            A2B = self.tm.get_transform(frame, "origin")
            collider = self.colliders_[frame]
            collider.update_pose(A2B)
            aabbs.append(collider.aabb())
        held = [d[0] for d in self.aabbtree_.external_data_list if d is not None]
        if held == frames and len(frames) > 0:
            self.aabbtree_.update_leaf_aabbs(aabbs)
            return
        self.aabbtree_ = AabbTree()
        for frame, aabb in zip(frames, aabbs):
            self.aabbtree_.insert_aabb(aabb, (frame, self.colliders_[frame]))

    def get_colliders(self):
        return self.colliders_.values()

    def get_artists(self):
        return [c.artist_ for c in self.colliders_.values() if c.artist_ is not None]

    def aabb_overlapping_colliders(self, collider, whitelist=()):
        """Frames -> colliders whose AABB overlaps ``collider``'s."""
        _, overlaps = self.aabbtree_.overlaps_aabb(collider.aabb())
        colliders = dict(self.aabbtree_.external_data_list[k] for k in overlaps)
        for frame in whitelist:
            colliders.pop(frame, None)
        return colliders

    def aabb_overlapping_with_other_bvh(self, other_bvh):
        _, _, _, pairs = self.aabbtree_.overlaps_aabb_tree(other_bvh.aabbtree_)
        return [(self.aabbtree_.external_data_list[a], other_bvh.aabbtree_.external_data_list[b])
                for a, b in pairs]

    def aabb_overlapping_with_self(self):
        _, _, _, pairs = self.aabbtree_.overlaps_aabb_tree(self.aabbtree_)
        return [(self.aabbtree_.external_data_list[a], self.aabbtree_.external_data_list[b])
                for a, b in pairs if a != b]

    def get_collider_frames(self):
        return self.collider_frames
