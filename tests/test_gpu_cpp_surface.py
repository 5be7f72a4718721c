"""The reference-named C++ surface (include/gwn/gwn.hpp) on the GPU: build (widths 2, 3, 4; LBVH
and H-PLOC), order-2 moments, exact and Taylor batches, then update + bounds refit + moment refit
+ Taylor again, all through the header's templates (tests/cpp/gwn_surface_gpu.cpp), compared with
the oracle: exact within the reference's 2e-5 + 2e-5 |w| of the FP64 winding, Taylor within 1e-5
of the oracle's replay on the same (and the refit) tree, and the tree's shape equal to the
oracle's."""

import ctypes
import os

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libgwn_surface_gpu.so")


@pytest.fixture(scope="module")
def surface(gwn):
    if not os.path.exists(SO):
        pytest.fail(f"{SO} is not built (tests/cpp/build.py)")
    lib = ctypes.CDLL(SO)
    lib.gwn_surface_workflow.restype = ctypes.c_int
    lib.gwn_surface_message.restype = ctypes.c_char_p
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("width", [2, 3, 4])
@pytest.mark.parametrize("method", ["hploc", "lbvh"])
def test_cpp_surface_workflow_matches_oracle(surface, orc, width, method):
    print(f"[cpp surface] W={width} {method}", flush=True)
    mesh = synth.displaced_sphere(64, 32)
    rng = np.random.default_rng(9)
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1)
    mv = (v * np.float32(1.1) + rng.normal(scale=0.005, size=v.shape).astype(np.float32)).astype(np.float32)
    moved = synth._mesh_from_arrays(mv, np.stack([mesh.i0, mesh.i1, mesh.i2], 1))
    q = synth.uniform_box_points(13, 8192, (-1.5,) * 3, (1.5,) * 3)
    n = q[0].shape[0]
    exact, taylor, refit = (np.zeros(n, np.float32) for _ in range(3))
    info = np.zeros(4, np.uint64)
    rc = surface.gwn_surface_workflow(
        _p(mesh.vx), _p(mesh.vy), _p(mesh.vz), ctypes.c_uint64(mesh.vertex_count), _p(mesh.i0), _p(mesh.i1),
        _p(mesh.i2), ctypes.c_uint64(mesh.triangle_count), _p(moved.vx), _p(moved.vy), _p(moved.vz),
        _p(q[0]), _p(q[1]), _p(q[2]), ctypes.c_uint64(n), ctypes.c_int(width),
        ctypes.c_int(0 if method == "hploc" else 1), _p(exact), _p(taylor), _p(refit), _p(info))
    assert rc == 0, surface.gwn_surface_message().decode()
    ref = orc.build_bvh4(mesh, method, 8, width=width)
    assert (int(info[0]), int(info[1]), int(info[2]), int(info[3])) == (
        ref.node_count, ref.internal_stack_bound, ref.packed_stack_bound, width)
    w64 = orc.winding_exact_mesh_f64(mesh, *q)
    assert np.all(np.abs(exact - w64) <= 2e-5 + 2e-5 * np.abs(w64))
    want = orc.winding_taylor_f32(ref, orc.refit_moment(ref, 2), 2, *q, accuracy_scale=2.0, stack_capacity=128)
    assert (np.abs(taylor.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 1e-5
    rref = orc.refit_bvh(ref, moved)
    want = orc.winding_taylor_f32(rref, orc.refit_moment(rref, 2), 2, *q, accuracy_scale=2.0, stack_capacity=128)
    assert (np.abs(refit.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 1e-5
