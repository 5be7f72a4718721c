"""The C-ABI library loads, exports exactly what include/gwn_hip.h declares, and its host-side
argument validation answers with the reference's diagnostics -- no GPU needed (stream 0,
no kernel launched)."""

import ctypes
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "smallgwn_amd", "libgwn_hip.so")
HEADER = os.path.join(REPO, "include", "gwn_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gwn_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libgwn_hip.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_header_declarations_are_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_python_mirror_lists_every_export():
    from smallgwn_amd import gwn

    assert sorted(gwn.EXPORTED_SYMBOLS) == declared_functions()


def test_abi_version(lib):
    assert lib.gwn_abi_version() == 3


def _err(lib):
    lib.gwn_last_error_message.restype = ctypes.c_char_p
    return lib.gwn_last_error_message().decode()


def test_host_validation_paths(lib):
    vp = ctypes.c_void_p
    g, b, m = vp(), vp(), vp()
    assert lib.gwn_geometry_create(ctypes.byref(g)) == 0
    assert lib.gwn_bvh_create(ctypes.byref(b)) == 0
    # Invalid build method and H-PLOC radius (include/gwn/detail/gwn_bvh_build_impl.cuh:62-72).
    lib.gwn_build_bvh4_f32.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_uint32, vp]
    assert lib.gwn_build_bvh4_f32(g, b, 7, 8, None) == 1
    assert "BVH build method is invalid" in _err(lib)
    assert lib.gwn_build_bvh4_f32(g, b, 0, 0, None) == 1
    assert "H-PLOC search radius must be in the range [1, 8]" in _err(lib)
    assert lib.gwn_build_bvh4_f32(g, b, 0, 9, None) == 1
    # Moment orders outside {0, 1, 2}.
    lib.gwn_moment_create.argtypes = [vp, ctypes.c_int32]
    assert lib.gwn_moment_create(ctypes.byref(m), 3) == 1
    assert "Taylor order must be 0, 1, or 2" in _err(lib)
    # Unbuilt BVH rejected before any launch (include/gwn/gwn_query.cuh:324-325).
    lib.gwn_compute_winding_number_exact_batch_f32.argtypes = [vp] * 5 + [ctypes.c_uint64, vp, vp]
    assert lib.gwn_compute_winding_number_exact_batch_f32(b, None, None, None, None, 0, None, None) == 1
    assert "BVH object contains no queryable data" in _err(lib)
    # Upload validation (include/gwn/detail/gwn_geometry_impl.cuh:80-130).
    lib.gwn_upload_geometry_f32.argtypes = [vp, vp, vp, vp, ctypes.c_uint64, vp, vp, vp, ctypes.c_uint64, vp]
    x = np.zeros(3, np.float32)
    idx = np.array([0, 1, 5], np.uint32)
    assert lib.gwn_upload_geometry_f32(g, x.ctypes.data, x.ctypes.data, x.ctypes.data, 3,
                                       idx.ctypes.data, idx.ctypes.data, idx.ctypes.data, 3, None) == 1
    assert "Triangle indices must be in [0, vertex_count)" in _err(lib)
    assert lib.gwn_upload_geometry_f32(g, None, None, None, 3, None, None, None, 0, None) == 1
    assert "non-null storage" in _err(lib)
    lib.gwn_bvh_destroy.argtypes = [vp]
    lib.gwn_geometry_destroy.argtypes = [vp]
    assert lib.gwn_bvh_destroy(b) == 0
    assert lib.gwn_geometry_destroy(g) == 0


def test_sort_argument_validation(lib):
    vp = ctypes.c_void_p
    lib.gwn_radix_sort_pairs_u64.argtypes = [vp, vp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, vp]
    assert lib.gwn_radix_sort_pairs_u64(None, None, 4, 0, 64, None) == 1
    assert lib.gwn_radix_sort_pairs_u64(None, None, 0, 0, 65, None) == 1
    assert "bit range" in _err(lib)
