"""Real = double and Index = uint64 objects on the GPU, against the oracle's double functions.

- The reference's FP64 builder workflow (tests/integration_gwn_builder.cu:345-435): 33 identical
  triangles with double vertices and uint64 indices, an H-PLOC build with radius 3, order-2
  moments, then the exact and the Taylor batch (accuracy 1e6, so every node is descended) within
  1e-12 of the FP64 exact sum, through the C++ surface (gwn.hpp) and through the Python binding.
- A displaced sphere whose vertices are not float values: the tree's shape equals the oracle's
  float build over the rounded positions, every double record and child box equals the oracle's
  double tree exactly, the double moments match the oracle's within 1e-12 of their scale, the exact
  batch is within 1e-12 of the FP64 sum and the Taylor batch (orders 0-2) within 1e-10 of the
  oracle's double traversal; then the same after an update and refits.
- uint64 indices with float positions give the float path's results bit for bit, and the f32 and
  f64 entry points refuse each other's objects.
"""

import ctypes
import os

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libgwn_surface_gpu.so")
CHILD_D = np.dtype([("bounds", "<f8", (6,)), ("reference", "<u8")])


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dev(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def _perturbed(mesh, seed=11, rel=1e-9):
    v = np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)
    u = synth.uniform01(seed, v.size).reshape(v.shape) * 2.0 - 1.0
    return v * (1.0 + rel * u)


def _nodes_d(raw, width):
    """Downloaded gwn_bvh_node<Width, double> bytes as (N, 4) children, empty slots past Width."""
    ch = np.frombuffer(raw.tobytes(), dtype=CHILD_D).reshape(-1, width)
    out = np.zeros((ch.shape[0], 4), dtype=CHILD_D)
    out[:, :width] = ch
    return out


# The reference test's inputs (integration_gwn_builder.cu:348-382).
REF_X, REF_Y, REF_Z = [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]
REF_Q = ([0.2, 0.7, -0.3], [0.2, 0.1, 0.3], [1.0, -2.0, 0.5])
REF_T = 33


def test_fp64_uint64_hploc_workflow_through_cpp_surface(orc):
    if not os.path.exists(SO):
        pytest.fail(f"{SO} is not built (tests/cpp/build.py)")
    print("[double] reference FP64 workflow via gwn.hpp", flush=True)
    lib = ctypes.CDLL(SO)
    lib.gwn_surface_workflow_f64_u64.restype = ctypes.c_int
    lib.gwn_surface_message.restype = ctypes.c_char_p
    vx, vy, vz = (np.array(a, np.float64) for a in (REF_X, REF_Y, REF_Z))
    i0, i1, i2 = (np.full(REF_T, k, np.uint64) for k in (0, 1, 2))
    q = [np.array(a, np.float64) for a in REF_Q]
    exact, taylor, refit = (np.zeros(3, np.float64) for _ in range(3))
    info = np.zeros(4, np.uint64)
    # The moved positions of the dynamic half: the same triangle lifted off z = 0.
    mz = np.array([0.25, 0.25, 0.25], np.float64)
    rc = lib.gwn_surface_workflow_f64_u64(
        _p(vx), _p(vy), _p(vz), ctypes.c_uint64(3), _p(i0), _p(i1), _p(i2), ctypes.c_uint64(REF_T),
        _p(vx), _p(vy), _p(mz), _p(q[0]), _p(q[1]), _p(q[2]), ctypes.c_uint64(3), ctypes.c_int(0),
        ctypes.c_uint32(3), ctypes.c_double(1e6), _p(exact), _p(taylor), _p(refit), _p(info))
    assert rc == 0, lib.gwn_surface_message().decode()
    ref = orc.winding_exact_mesh_d64(np.stack([vx, vy, vz]), i0, i1, i2, *q)
    assert np.abs(exact - ref).max() <= 1e-12
    assert np.abs(taylor - ref).max() <= 1e-12
    moved = orc.winding_exact_mesh_d64(np.stack([vx, vy, mz]), i0, i1, i2, *q)
    assert np.abs(refit - moved).max() <= 1e-12
    assert int(info[3]) == 4


def test_fp64_uint64_hploc_workflow_through_python(gwn, orc, dev):
    import torch

    print("[double] reference FP64 workflow via the Python binding", flush=True)
    g = gwn.GeometryObject()
    i0, i1, i2 = (np.full(REF_T, k, np.uint64) for k in (0, 1, 2))
    gwn.gwn_upload_geometry(g, REF_X, REF_Y, REF_Z, i0, i1, i2, real=np.float64, index=np.uint64).raise_if_error()
    b = gwn.BvhObject()
    gwn.gwn_build_bvh(g, b, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=3)).raise_if_error()
    info = b.info()
    raw, prim, rec = b.download()
    nodes = _nodes_d(raw, info["width"])
    # Every internal node is complete (integration_gwn_builder.cu:36-119): 2-4 valid children, each
    # triangle reached once, boxes = the triangle's box.
    seen = np.zeros(REF_T, int)
    for n in range(info["node_count"]):
        refs = [int(c["reference"]) for c in nodes[n]]
        valid = [r for r in refs if r >> 47 & 1]
        assert 2 <= len(valid) <= 4
        for c in nodes[n]:
            r = int(c["reference"])
            if r >> 47 & 1 and r >> 48:
                seen[r & ((1 << 47) - 1)] += 1
                assert np.array_equal(c["bounds"], [0.0, 0.0, 0.0, 1.0, 1.0, 0.0])
    assert np.all(seen == 1) and sorted(prim.tolist()) == list(range(REF_T))
    assert np.array_equal(rec, np.tile([0.0, 0, 0, 1, 0, 0, 0, 1, 0], (REF_T, 1)))
    m = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(b, m).raise_if_error()
    q = [_dev(a, dev) for a in REF_Q]
    ex, ta = torch.empty(3, dtype=torch.float64, device=dev), torch.empty(3, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(b, *q, ex).raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(b, m, *q, ta, accuracy_scale=1e6).raise_if_error()
    torch.cuda.synchronize()
    ref = orc.winding_exact_mesh_d64(np.array([REF_X, REF_Y, REF_Z]), i0, i1, i2, *REF_Q)
    assert np.abs(ex.cpu().numpy() - ref).max() <= 1e-12
    assert np.abs(ta.cpu().numpy() - ref).max() <= 1e-12


@pytest.mark.parametrize("method,radius", [("hploc", 8), ("lbvh", 8)])
def test_fp64_displaced_sphere_matches_double_oracle(gwn, orc, dev, method, radius):
    import torch

    print(f"[double] displaced sphere {method}", flush=True)
    mesh = synth.displaced_sphere(48, 24)
    vd = _perturbed(mesh)
    assert not np.array_equal(vd, vd.astype(np.float32).astype(np.float64)), "vertices must not be float values"
    rounded = synth._mesh_from_arrays(vd.T.astype(np.float32), np.stack([mesh.i0, mesh.i1, mesh.i2], 1))
    g = gwn.GeometryObject()
    gwn.gwn_upload_geometry(g, vd[0], vd[1], vd[2], mesh.i0, mesh.i1, mesh.i2, real=np.float64).raise_if_error()
    b = gwn.BvhObject()
    opts = gwn.BuildOptions(method=gwn.BuildMethod.HPLOC if method == "hploc" else gwn.BuildMethod.LBVH,
                            hploc_search_radius=radius)
    gwn.gwn_build_bvh(g, b, opts).raise_if_error()
    ref = orc.build_bvh4(rounded, method, radius)
    info = b.info()
    assert (info["node_count"], info["internal_stack_bound"], info["packed_stack_bound"]) == (
        ref.node_count, ref.internal_stack_bound, ref.packed_stack_bound)
    raw, prim, rec = b.download()
    nodes = _nodes_d(raw, info["width"])
    assert np.array_equal(prim, ref.primitive_indices)
    assert np.array_equal(nodes["reference"], np.stack([[c["reference"] for c in n["children"]] for n in ref.nodes]))
    rec_o, child_o, _ = orc.double_tree(ref, vd, mesh.i0, mesh.i1, mesh.i2)
    assert np.array_equal(rec, rec_o)
    assert np.array_equal(nodes["bounds"], child_o)

    q = synth.uniform_box_points(21, 2048, (-1.5,) * 3, (1.5,) * 3)
    qd = [a.astype(np.float64) + 1e-10 for a in q]  # double queries, not float values either
    tq = [_dev(a, dev) for a in qd]
    out = torch.empty(len(qd[0]), dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(b, *tq, out).raise_if_error()
    torch.cuda.synchronize()
    exact = orc.winding_exact_mesh_d64(vd, mesh.i0, mesh.i1, mesh.i2, *qd)
    assert np.abs(out.cpu().numpy() - exact).max() <= 1e-12

    for order in (0, 1, 2):
        m = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(b, m).raise_if_error()
        got_m = m.download_f64(info["node_count"], info["width"])
        want_m = orc.refit_moment_d64(ref, rec_o, child_o, order)
        scale = np.abs(want_m).max(axis=(0, 2), keepdims=True) + 1e-300
        assert (np.abs(got_m - want_m) / scale).max() <= 1e-12, order
        gwn.gwn_compute_winding_number_taylor_batch(b, m, *tq, out, accuracy_scale=2.0).raise_if_error()
        torch.cuda.synchronize()
        want = orc.winding_taylor_d64(ref, rec_o, want_m, order, *qd, accuracy_scale=2.0, stack_capacity=128)
        assert np.abs(out.cpu().numpy() - want).max() <= 1e-10, order

    # Dynamic: new double positions, bounds and moment refits, query again.
    moved = vd * 1.05 + 1e-3
    gwn.gwn_update_geometry(g, moved[0], moved[1], moved[2]).raise_if_error()
    gwn.gwn_refit_bvh(g, b).raise_if_error()
    raw, _, rec = b.download()
    rec_m, child_m, _ = orc.double_tree(ref, moved, mesh.i0, mesh.i1, mesh.i2)
    assert np.array_equal(rec, rec_m) and np.array_equal(_nodes_d(raw, info["width"])["bounds"], child_m)
    m = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(b, m).raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(b, m, *tq, out, accuracy_scale=2.0).raise_if_error()
    torch.cuda.synchronize()
    want = orc.winding_taylor_d64(ref, rec_m, orc.refit_moment_d64(ref, rec_m, child_m, 2), 2, *qd,
                                  accuracy_scale=2.0, stack_capacity=128)
    assert np.abs(out.cpu().numpy() - want).max() <= 1e-10


def test_uint64_indices_with_float_positions_equal_uint32(gwn, dev):
    import torch

    mesh = synth.torus(32, 16)
    q = [torch.from_numpy(a).to(dev) for a in synth.uniform_box_points(4, 1024, (-1.5,) * 3, (1.5,) * 3)]
    outs = []
    for index in (np.uint32, np.uint64):
        g = gwn.GeometryObject()
        gwn.gwn_upload_geometry(g, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2, index=index).raise_if_error()
        b = gwn.BvhObject()
        gwn.gwn_build_bvh(g, b).raise_if_error()
        m = gwn.MomentObject(1)
        gwn.gwn_refit_bvh_moment(b, m).raise_if_error()
        o = torch.empty(1024, device=dev)
        gwn.gwn_compute_winding_number_taylor_batch(b, m, *q, o).raise_if_error()
        torch.cuda.synchronize()
        outs.append(o.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    g = gwn.GeometryObject()
    bad = np.array([0, 1, 2**40], np.uint64)
    st = gwn.gwn_upload_geometry(g, mesh.vx, mesh.vy, mesh.vz, bad, bad, bad, index=np.uint64)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "vertex_count" in st.message


def test_float_and_double_entry_points_refuse_each_other(gwn, dev):
    import torch

    mesh = synth.icosphere(2)
    g = gwn.GeometryObject()
    gwn.gwn_upload_geometry(g, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2, real=np.float64).raise_if_error()
    b = gwn.BvhObject()
    gwn.gwn_build_bvh(g, b).raise_if_error()
    lib = gwn.library()
    q = torch.zeros(4, device=dev)
    # Float queries on a double BVH, a float build of a double geometry, a float moment refit.
    assert lib.gwn_compute_unsigned_distance_batch_f32(b.handle, q.data_ptr(), q.data_ptr(), q.data_ptr(),
                                                       q.data_ptr(), 4, float("inf"), None, None) == 1
    assert "f64" in gwn.library().gwn_last_error_message().decode()
    b32 = gwn.BvhObject()
    assert lib.gwn_build_bvh_f32(g.handle, b32.handle, 4, 0, 8, None) == 1
    m = gwn.MomentObject(1)
    assert lib.gwn_refit_bvh_moment_f32(b.handle, m.handle, None) == 1
    # The binding refuses float32 spans for a double BVH before any launch.
    st = gwn.gwn_compute_winding_number_exact_batch(b, q, q, q, q)
    assert st.error == gwn.Error.INVALID_ARGUMENT
    # And the double entry points refuse a float geometry.
    g32 = gwn.GeometryObject()
    gwn.gwn_upload_geometry(g32, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    assert lib.gwn_build_bvh_f64(g32.handle, b32.handle, 4, 0, 8, None) == 1
    with pytest.raises(gwn.GwnError):
        b.accessor()
