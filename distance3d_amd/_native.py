"""ctypes binding of libd3d_hip.so (the C ABI declared in include/d3d.h).

The library is the only compute path of this package: every public function
that computes a support point, an AABB, a GJK/EPA result or a broad-phase pair
set calls into it.  There is no CPU fallback; when the library or a GPU is
missing, :func:`lib` raises.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libd3d_hip.so")

# collider types (include/d3d.h)
HULL, SPHERE, CAPSULE, ELLIPSOID, CYLINDER, BOX, CONE, DISK, ELLIPSE = range(9)

# GJK status codes
SEPARATED = 0
INTERSECT = 1
CLIPPED = 3
ITER_CAP = 4
SANITY_FAILED = 0x100
BAD_INPUT = 0x200


class D3DError(RuntimeError):
    """A libd3d_hip entry point returned an error code."""


class d3d_scene(ctypes.Structure):
    _fields_ = [
        ("shape_f64", ctypes.c_void_p),
        ("shape_i32", ctypes.c_void_p),
        ("vx", ctypes.c_void_p),
        ("vy", ctypes.c_void_p),
        ("vz", ctypes.c_void_p),
        ("v64", ctypes.c_void_p),
        ("n_shapes", ctypes.c_int64),
        ("n_vertices", ctypes.c_int64),
        ("n_vertices64", ctypes.c_int64),
        ("vorig", ctypes.c_void_p),
        ("clusters", ctypes.c_void_p),
        ("n_clusters", ctypes.c_int64),
    ]


ABI_VERSION = 5
CLUSTER_SIZE = 16
NOALLOC_MAX_PAIRS = 16384  # D3D_NOALLOC_MAX_PAIRS


class d3d_gjk_params(ctypes.Structure):
    _fields_ = [
        ("tolerance", ctypes.c_double),
        ("max_distance_squared", ctypes.c_double),
        ("sanity_check", ctypes.c_double),
        ("max_iterations", ctypes.c_int32),
        ("hull_serial_max", ctypes.c_int32),
        ("lanes_per_pair", ctypes.c_int32 * 16),
        ("blocks_per_cu", ctypes.c_int32),
        ("claim_factor", ctypes.c_int32),
        ("epa_blocks_per_cu", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double
_SCENE = ctypes.POINTER(d3d_scene)

# name -> argtypes; every function returns int
_SIGNATURES = {
    "d3d_support_batch": [_SCENE, _P, _P, _I64, _P, _P],
    "d3d_aabb_batch": [_SCENE, _P, _I64, _P, _P],
    "d3d_gjk_distance_batch": [_SCENE, _P, _P, _I64, ctypes.POINTER(d3d_gjk_params),
                               _P, _P, _P, _P, _P, _P, _P, _P],
    "d3d_gjk_distance_batch_ws": [_SCENE, _P, _P, _I64, ctypes.POINTER(d3d_gjk_params),
                                  _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P],
    "d3d_gjk_intersection_batch": [_SCENE, _P, _P, _I64, _D, _I32, _P, _P, _P],
    "d3d_epa_batch": [_SCENE, _P, _P, _I64, _P, _P, _I32, _I32, _D, _P, _P, _P, _P, _P, _P],
    "d3d_gjk_epa_batch_ws": [_SCENE, _P, _P, _I64, ctypes.POINTER(d3d_gjk_params), _I32, _I32,
                             _D, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P],
    "d3d_gjk_epa_batch_noalloc": [_SCENE, _P, _P, _I64, ctypes.POINTER(d3d_gjk_params), _I32,
                                  _I32, _D, _P, _P, _P, _P, _P, _P, _P, _P],
    "d3d_stage_timing": [_I32],
    "d3d_stage_timing_read": [_P, _P],
    "d3d_instance_pairs": [_SCENE, _P, _P, _P, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P],
    "d3d_bvh_build": [_P, _I64, _P, _I64, _P],
    "d3d_bvh_query": [_P, _I64, _P, _P, _I64, _P, _I64, _P, _P],
    "d3d_bvh_refit": [_P, _I64, _P, _I64, _P],
    "d3d_overlap_dense": [_P, _I64, _P, _I64, _P, _I64, _P, _P, _I64, _P],
    "d3d_mpr_batch": [_P, _P, _P, _P, _I64, _D, _I32, _I32, _P, _P, _P],
    "d3d_tet_pairs_batch": [_P, _P, _P, _P, _P, _P, _P, _I64, _D, _D, _P, _P, _P, _P, _P, _I64,
                            _P],
    "d3d_tet_mesh_prep": [_P, _I64, _P, _P, _P, _P],
    "d3d_polygon_angles": [_P, _P, _I64, _P, _P],
    "d3d_intersect_halfplanes": [_P, _I32, _P, _P, _P],
    "d3d_tet_contact_forces": [_P, _P, _P, _P, _P, _P, _P, _I64, _D, _P, _P],
    "d3d_cluster_plan": [_P, _I64, _I32, _P, _P, _P],
    "d3d_cluster_scene": [_P, _P, _I64, _P, _P, _P, _P, _I64, _I64, _I32, _I32,
                          _P, _P, _P, _P, _P, _P, _P, _P],
}

_lock = threading.Lock()
_lib = None
_initialised = False


def exported_symbols():
    """Names of the C-ABI functions this binding uses."""
    return (["d3d_abi_version", "d3d_last_error", "d3d_init", "d3d_gjk_workspace_bytes",
             "d3d_bvh_bytes", "d3d_gjk_epa_workspace_bytes", "d3d_tet_pairs_workspace_bytes",
             "d3d_overlap_dense_workspace_bytes"]
            + list(_SIGNATURES))


def load_library(path=LIB_PATH):
    """Load libd3d_hip.so without touching the GPU (no HIP call is made)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise D3DError(
                    "%s is missing: build it with __graft_entry__.build() or "
                    "`make -C distance3d_amd/csrc`" % path)
            lib = ctypes.CDLL(path)
            lib.d3d_abi_version.restype = ctypes.c_int
            if lib.d3d_abi_version() != ABI_VERSION:
                raise D3DError("%s has ABI version %d, this binding needs %d: rebuild it"
                               % (path, lib.d3d_abi_version(), ABI_VERSION))
            lib.d3d_last_error.restype = ctypes.c_char_p
            lib.d3d_init.argtypes = [ctypes.c_int]
            lib.d3d_init.restype = ctypes.c_int
            for name in ("d3d_gjk_workspace_bytes", "d3d_bvh_bytes",
                         "d3d_gjk_epa_workspace_bytes", "d3d_tet_pairs_workspace_bytes",
                         "d3d_overlap_dense_workspace_bytes"):
                getattr(lib, name).argtypes = [ctypes.c_int64]
                getattr(lib, name).restype = ctypes.c_int64
            for name, argtypes in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = argtypes
                fn.restype = ctypes.c_int
            _lib = lib
        return _lib


def lib():
    """The loaded library, initialised on the current torch CUDA device."""
    global _initialised
    import torch
    l = load_library()
    if not _initialised:
        if not torch.cuda.is_available():
            raise D3DError("distance3d_amd needs a ROCm GPU (gfx950); none is visible")
        check(l.d3d_init(torch.cuda.current_device()), "d3d_init")
        _initialised = True
    return l


def check(rc, what):
    if rc != 0:
        msg = _lib.d3d_last_error().decode() if _lib is not None else ""
        raise D3DError("%s failed (code %d): %s" % (what, rc, msg))


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


# work classes: size class 0..7 (total hull vertices 0, <=32, <=128, <=512,
# <=1024, <=2048, <=4096, more), plus 8 when both colliders are hulls
NUM_CLASSES = 16
DEFAULT_LANES = (1, 4, 16, 8, 16, 16, 32, 64, 1, 8, 16, 8, 64, 16, 64, 64)


def expand_lanes(lanes):
    """A lanes table for the 16 work classes: 16 entries as given, or 8 (one
    per size class) used for pairs of one hull and of two hulls alike."""
    lanes = [int(g) for g in lanes]
    if len(lanes) == NUM_CLASSES // 2:
        lanes = lanes * 2
    if len(lanes) != NUM_CLASSES:
        raise ValueError("lanes_per_pair needs 16 entries (one per work class) or 8 (one per "
                         "size class)")
    return lanes


def gjk_params(tolerance=1e-10, max_distance_squared=100000.0, sanity_check=1e-8,
               max_iterations=128, hull_serial_max=16, lanes_per_pair=None, launch=None):
    """d3d_gjk_params.  ``launch``: optional dict of the launch shape
    (``blocks_per_cu``, ``claim_factor``, ``epa_blocks_per_cu``; 0 or absent =
    the library default)."""
    p = d3d_gjk_params()
    p.tolerance = float(tolerance)
    p.max_distance_squared = float(max_distance_squared)
    p.sanity_check = float(sanity_check)
    p.max_iterations = int(max_iterations)
    p.hull_serial_max = int(hull_serial_max)
    lanes = expand_lanes(DEFAULT_LANES if lanes_per_pair is None else lanes_per_pair)
    for c, g in enumerate(lanes):
        p.lanes_per_pair[c] = int(g)
    launch = launch or {}
    unknown = set(launch) - {"blocks_per_cu", "claim_factor", "epa_blocks_per_cu"}
    if unknown:
        raise ValueError("unknown launch keys: %s" % sorted(unknown))
    p.blocks_per_cu = int(launch.get("blocks_per_cu", 0))
    p.claim_factor = int(launch.get("claim_factor", 0))
    p.epa_blocks_per_cu = int(launch.get("epa_blocks_per_cu", 0))
    return p


EPSILON = float(np.finfo(float).eps)
MAX_FLOAT = float(np.finfo(float).max)
