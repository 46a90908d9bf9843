"""The C-ABI library loads and exports every symbol include/d3d.h declares
(no GPU call is made)."""
import os
import re

import pytest

from distance3d_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^(?:int|int64_t|const char \*)\s*(d3d_\w+)\(", src, re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "d3d_gjk_distance_batch" in names
    assert "d3d_last_error" in names


def test_library_exports_declared_symbols():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libd3d_hip.so not built; run __graft_entry__.build()")
    lib = _native.load_library()
    for name in sorted(_declared()):
        assert hasattr(lib, name), name
    assert lib.d3d_abi_version() == _native.ABI_VERSION


def test_binding_covers_header():
    assert _declared() == set(_native.exported_symbols())


def test_argument_errors_without_gpu():
    """Bad arguments are rejected with D3D_E_INVALID and a message before any
    HIP call, and empty inputs return D3D_OK without one (runs with no GPU)."""
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libd3d_hip.so not built; run __graft_entry__.build()")
    lib = _native.load_library()
    invalid = 1
    assert lib.d3d_tet_mesh_prep(None, 5, None, None, None, None) == invalid
    assert b"NULL" in lib.d3d_last_error()
    assert lib.d3d_tet_mesh_prep(None, -1, None, None, None, None) == invalid
    assert lib.d3d_tet_mesh_prep(None, 0, None, None, None, None) == 0
    assert lib.d3d_tet_contact_forces(None, None, None, None, None, None, None, 4, 1.0, None,
                                      None) == invalid
    assert lib.d3d_tet_contact_forces(None, None, None, None, None, None, None, 0, 1.0, None,
                                      None) == 0
    assert lib.d3d_polygon_angles(None, None, 3, None, None) == invalid
    assert lib.d3d_polygon_angles(None, None, 0, None, None) == 0
    assert lib.d3d_overlap_dense(None, 3, None, 3, None, 0, None, None, 0, None) == invalid
    assert lib.d3d_overlap_dense(None, -1, None, 3, None, 0, None, None, 0, None) == invalid
    # workspace sizes: pure host arithmetic
    assert lib.d3d_overlap_dense_workspace_bytes(-1) == -1
    assert lib.d3d_overlap_dense_workspace_bytes(5) == 48
    assert lib.d3d_tet_pairs_workspace_bytes(-1) == -1
    assert lib.d3d_tet_pairs_workspace_bytes(0) == 0
    assert lib.d3d_tet_pairs_workspace_bytes(100) == 128 * 48 * 8
    assert lib.d3d_tet_pairs_workspace_bytes(1 << 30) == 4096 * 64 * 48 * 8
    # intersect_halfplanes: one wavefront does O(n^3) work; n is capped
    assert lib.d3d_intersect_halfplanes(None, 1025, None, None, None) == invalid
    assert b"1024" in lib.d3d_last_error()
    assert lib.d3d_intersect_halfplanes(None, -1, None, None, None) == invalid
    # EPA: max_iterations below 1 and max_faces outside [4, 128] are rejected
    # before any pointer is used (the pointers below are never dereferenced)
    import ctypes
    sc = _native.d3d_scene()
    sc.shape_f64 = sc.shape_i32 = 16
    sc.n_shapes = 1
    fake = ctypes.c_void_p(16)
    args = [ctypes.byref(sc), fake, fake, 1, fake, fake]
    # scene arrays read with 16-byte loads must be 16-byte aligned
    sc.shape_i32 = 8
    assert lib.d3d_epa_batch(*args, 64, 128, 1e-8, fake, fake, None, None, None, None) == invalid
    assert b"aligned" in lib.d3d_last_error()
    sc.shape_i32 = 16
    assert lib.d3d_epa_batch(*args, 0, 64, 1e-8, fake, fake, None, None, None, None) == invalid
    assert b"max_iterations" in lib.d3d_last_error()
    assert lib.d3d_epa_batch(*args, 64, 129, 1e-8, fake, fake, None, None, None, None) == invalid
    assert b"max_faces" in lib.d3d_last_error()
    assert lib.d3d_epa_batch(*args, 64, 3, 1e-8, fake, fake, None, None, None, None) == invalid
    # allocation-free pipeline: sizes beyond its array are rejected, n == 0 is
    # a no-op, both before any HIP call
    big = _native.NOALLOC_MAX_PAIRS + 1
    noalloc = [ctypes.byref(sc), fake, fake, big, None, 64, 128, 1e-8] + [fake] * 7 + [None]
    assert lib.d3d_gjk_epa_batch_noalloc(*noalloc) == invalid
    assert b"D3D_NOALLOC_MAX_PAIRS" in lib.d3d_last_error()
    noalloc[3] = -1
    assert lib.d3d_gjk_epa_batch_noalloc(*noalloc) == invalid
    noalloc[3] = 0
    assert lib.d3d_gjk_epa_batch_noalloc(*noalloc) == 0
