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
