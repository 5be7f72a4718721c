"""GPU parity of the build path: onesweep sort, Morton/LBVH/BVH4 build and Taylor-moment refit.

Integer and index results (sorted order, node arrays, primitive order, records, stack bounds)
must be bit-identical to the oracle; moments too, because both sides evaluate the same unfused
IEEE float sequence.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests import helpers

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("n", [0, 1, 2, 100, 4095, 4096, 4097, 100_003, 1 << 20])
def test_radix_sort_u64_stable(gwn, dev, n):
    torch = _torch()
    rng = np.random.default_rng(n)
    keys = (rng.integers(0, 1 << 20, size=n, dtype=np.uint64) << np.uint64(40)) | rng.integers(0, 7, size=n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    dk = torch.from_numpy(keys.view(np.int64)).to(dev)
    dv = torch.from_numpy(vals.view(np.int32)).to(dev)
    gwn.radix_sort_pairs(dk, dv, 0, 64).raise_if_error()
    torch.cuda.synchronize()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(dk.cpu().numpy().view(np.uint64), keys[order])
    assert np.array_equal(dv.cpu().numpy().view(np.uint32), vals[order])


@pytest.mark.parametrize("bits", [(0, 32), (8, 24), (0, 30)])
def test_radix_sort_u32_bit_ranges(gwn, dev, bits):
    torch = _torch()
    rng = np.random.default_rng(bits[0] + bits[1])
    n = 300_001
    keys = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    dk = torch.from_numpy(keys.view(np.int32)).to(dev)
    dv = torch.from_numpy(vals.view(np.int32)).to(dev)
    gwn.radix_sort_pairs(dk, dv, bits[0], bits[1]).raise_if_error()
    torch.cuda.synchronize()
    width = bits[1] - bits[0]
    sub = (keys >> np.uint32(bits[0])) & np.uint32((1 << width) - 1 if width < 32 else 0xFFFFFFFF)
    order = np.argsort(sub, kind="stable")
    assert np.array_equal(dv.cpu().numpy().view(np.uint32), vals[order])
    assert np.array_equal(dk.cpu().numpy().view(np.uint32), keys[order])


def _meshes():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(3000, 3)).astype(np.float32)
    soup = synth._mesh_from_arrays(v, rng.integers(0, 3000, size=(2000, 3)))
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    return {
        "single": synth._mesh_from_arrays(tri, [[0, 1, 2]]),
        "two": synth._mesh_from_arrays(
            np.stack([[0, 1, 0, 4, 6, 4], [0, 0, 2, 0, 0, 3], [1, 1, 1, -2, -2, -2]], 1), [[0, 1, 2], [3, 4, 5]]),
        "asym_octahedron": synth.asymmetric_octahedron(),
        "coincident33": synth._mesh_from_arrays(tri, np.tile([[0, 1, 2]], (33, 1))),
        "icosphere4": synth.icosphere(4),
        "torus": synth.torus(64, 32),
        "soup": soup,
        "displaced_256x128": synth.displaced_sphere(256, 128),
    }


MESHES = _meshes()


def build_gpu(gwn, mesh, method=None):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method or gwn.BuildMethod.LBVH)).raise_if_error()
    return geom, bvh


def assert_bvh_matches_oracle(gwn, orc, mesh, bvh, method="lbvh", radius=8):
    ref = orc.build_bvh4(mesh, method=method, radius=radius)
    info = bvh.info()
    nodes, prim, tris = bvh.download()
    assert info["node_count"] == ref.node_count
    assert info["root_reference"] == ref.root_reference
    assert np.array_equal(info["root_bounds"].view(np.uint32), ref.root_bounds.view(np.uint32))
    assert info["internal_stack_bound"] == ref.internal_stack_bound
    assert info["packed_stack_bound"] == ref.packed_stack_bound
    assert np.array_equal(prim, ref.primitive_indices)
    assert np.array_equal(tris.view(np.uint32), ref.triangles.view(np.uint32))
    assert nodes.tobytes() == ref.nodes.tobytes()
    return ref


@pytest.mark.parametrize("name", list(MESHES))
def test_build_matches_oracle(gwn, orc, name):
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    ref = assert_bvh_matches_oracle(gwn, orc, mesh, bvh)
    nodes, prim, tris = bvh.download()
    helpers.check_bvh4(nodes, bvh.info()["root_reference"], prim, tris, mesh,
                       ref.internal_stack_bound, ref.packed_stack_bound)


def test_build_large_mesh_matches_oracle(gwn, orc):
    """The 1,048,576-triangle displaced sphere: node arrays bit-identical at full size."""
    mesh = synth.config_large_mesh()
    _, bvh = build_gpu(gwn, mesh)
    assert_bvh_matches_oracle(gwn, orc, mesh, bvh)


def test_rebuild_is_deterministic_and_revisions_advance(gwn):
    mesh = MESHES["torus"]
    geom, bvh = build_gpu(gwn, mesh)
    a = bvh.download()
    r1 = bvh.info()["revision"]
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    b = bvh.download()
    assert bvh.info()["revision"] > r1
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_empty_geometry_builds_empty_bvh(gwn):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, [], [], [], [], [], []).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    assert not bvh.has_data()


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("name", ["two", "asym_octahedron", "coincident33", "icosphere4", "soup", "displaced_256x128"])
def test_moment_matches_oracle_bitwise(gwn, orc, name, order):
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    mom = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    n = bvh.info()["node_count"]
    got = mom.download(n)
    ref = orc.refit_moment(orc.build_bvh4(mesh), order)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_moment_large_mesh_matches_oracle_bitwise(gwn, orc):
    mesh = synth.config_large_mesh()
    _, bvh = build_gpu(gwn, mesh)
    ref_bvh = orc.build_bvh4(mesh)
    for order in (1, 2):
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        got = mom.download(bvh.info()["node_count"])
        assert np.array_equal(got.view(np.uint32), orc.refit_moment(ref_bvh, order).view(np.uint32))


def test_repeated_builds_and_refits_are_bit_identical(gwn, orc):
    """Many builds and order-2 refits of one mesh: every one equals the oracle bit for bit."""
    mesh = MESHES["displaced_256x128"]
    ref = orc.build_bvh4(mesh)
    ref_m = orc.refit_moment(ref, 2)
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    for rep in range(40):
        bvh = gwn.BvhObject()
        gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
        nodes, prim, tris = bvh.download()
        assert nodes.tobytes() == ref.nodes.tobytes(), f"build {rep}: nodes differ"
        assert np.array_equal(prim, ref.primitive_indices), f"build {rep}: primitive order differs"
        mom = gwn.MomentObject(2)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        got = mom.download(bvh.info()["node_count"])
        assert np.array_equal(got.view(np.uint32), ref_m.view(np.uint32)), f"refit {rep}: moments differ"


# ---------------------------------------------------------------------------------------------
# H-PLOC builder (gwn_build_bvh with method = H-PLOC)
# ---------------------------------------------------------------------------------------------

def _equal_triangles(n):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    return synth._mesh_from_arrays(v, np.tile([[0, 1, 2]], (n, 1)))


def _hploc(gwn, mesh, radius=8, geom=None):
    if geom is None:
        geom = gwn.GeometryObject()
        gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC,
                                                  hploc_search_radius=radius)).raise_if_error()
    return geom, bvh


@pytest.mark.parametrize("radius", [1, 4, 8])
@pytest.mark.parametrize("name", list(MESHES))
def test_hploc_build_matches_oracle(gwn, orc, name, radius):
    """Node arrays, primitive order, records, root and stack bounds equal the oracle's H-PLOC
    restatement bit for bit (deterministic node ids on both sides)."""
    print(f"[hploc] {name} radius={radius}", flush=True)
    mesh = MESHES[name]
    _, bvh = _hploc(gwn, mesh, radius)
    ref = assert_bvh_matches_oracle(gwn, orc, mesh, bvh, method="hploc", radius=radius)
    nodes, prim, tris = bvh.download()
    helpers.check_bvh4(nodes, bvh.info()["root_reference"], prim, tris, mesh,
                       ref.internal_stack_bound, ref.packed_stack_bound)


@pytest.mark.parametrize("count", [2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 129, 513])
def test_hploc_repeated_equal_bounds_match_oracle(gwn, orc, count):
    """tests/integration_gwn_builder.cu:301-343: identical triangles around the list threshold,
    the wave width and multi-block launches, for radius 1, 4 and 8."""
    mesh = _equal_triangles(count)
    for radius in (1, 4, 8):
        print(f"[hploc] equal count={count} radius={radius}", flush=True)
        _, bvh = _hploc(gwn, mesh, radius)
        assert_bvh_matches_oracle(gwn, orc, mesh, bvh, method="hploc", radius=radius)


def test_hploc_large_mesh_matches_oracle(gwn, orc):
    """The 1,048,576-triangle displaced sphere built with H-PLOC (radius 8)."""
    mesh = synth.config_large_mesh()
    print("[hploc] large mesh", flush=True)
    _, bvh = _hploc(gwn, mesh)
    assert_bvh_matches_oracle(gwn, orc, mesh, bvh, method="hploc", radius=8)


def test_hploc_repeated_builds_are_bit_identical(gwn, orc):
    mesh = MESHES["displaced_256x128"]
    ref = orc.build_bvh4(mesh, method="hploc")
    geom = None
    for rep in range(20):
        geom, bvh = _hploc(gwn, mesh, 8, geom)
        nodes, prim, _ = bvh.download()
        assert nodes.tobytes() == ref.nodes.tobytes(), f"build {rep}: nodes differ"
        assert np.array_equal(prim, ref.primitive_indices)


def test_hploc_radius_validation(gwn):
    mesh = _equal_triangles(8)
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    for radius in (0, 9):
        st = gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=radius))
        assert st.error == gwn.Error.INVALID_ARGUMENT and "radius" in st.message
    assert not bvh.has_data()


def _builder_queries(mesh):
    """tests/integration_gwn_builder.cu:120-166: the box centre, six points 1.3 extents out along
    the axes and two interior points."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    lo, hi = v.min(axis=0), v.max(axis=0)
    c = (lo + hi) * np.float32(0.5)
    e = np.maximum(hi - lo, np.float32(1e-2))
    pts = [c]
    for axis in range(3):
        for sign in (1, -1):
            p = c.copy()
            p[axis] = c[axis] + np.float32(sign * 1.3) * e[axis]
            pts.append(p)
    pts.append(c + np.float32([0.31, -0.27, 0.23]) * e)
    pts.append(c + np.float32([-0.19, 0.37, -0.29]) * e)
    q = np.stack(pts).astype(np.float32)
    return q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy()


@pytest.mark.parametrize("name", ["icosphere4", "torus", "displaced_256x128", "soup"])
def test_hploc_tree_serves_queries(gwn, orc, dev, name):
    """tests/integration_gwn_builder.cu:169-239: exact and fully descended (beta = 1e6) Taylor
    winding on the H-PLOC tree match the FP64 exact winding within 2e-5 + 2e-5 |w|."""
    import torch

    mesh = MESHES[name]
    _, bvh = _hploc(gwn, mesh)
    mom = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    qx, qy, qz = _builder_queries(mesh)
    d = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (qx, qy, qz)]
    out = torch.zeros(qx.shape[0], dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, out, accuracy_scale=1e6).raise_if_error()
    ex = torch.zeros_like(out)
    gwn.gwn_compute_winding_number_exact_batch(bvh, *d, ex).raise_if_error()
    torch.cuda.synchronize()
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    for got in (out.cpu().numpy(), ex.cpu().numpy()):
        assert np.all(np.abs(got - ref) <= 2e-5 + 2e-5 * np.abs(ref))


# ---------------------------------------------------------------------------------------------
# Bounds refit (gwn_refit_bvh)
# ---------------------------------------------------------------------------------------------

def _deformed(mesh, seed):
    from tests.test_oracle import _deform

    return _deform(mesh, seed)


def test_refit_bvh_kat(gwn, orc):
    """tests/unit_gwn_bvh.cu:220-287 through the C ABI: root (-7,-4,-6)-(11,9,3), primitive order
    preserved, records from the new positions."""
    x, y, z = [0, 1, 0, 4, 6, 4], [0, 0, 2, 0, 0, 3], [1, 1, 1, -2, -2, -2]
    faces = [[0, 1, 2], [3, 4, 5]]
    mesh = synth._mesh_from_arrays(np.stack([x, y, z], 1), faces)
    geom, bvh = build_gpu(gwn, mesh)
    _, prim_before, _ = bvh.download()
    moved = synth._mesh_from_arrays(
        np.stack([[-7, -2, -7, 8, 11, 8], [5, 5, 9, -4, -4, 2], [3, 3, 3, -6, -6, -6]], 1), faces)
    rev = bvh.info()["revision"]
    gwn.gwn_update_geometry(geom, moved.vx, moved.vy, moved.vz).raise_if_error()
    gwn.gwn_refit_bvh(geom, bvh).raise_if_error()
    info = bvh.info()
    assert info["root_bounds"].tolist() == [-7, -4, -6, 11, 9, 3]
    assert info["revision"] > rev
    nodes, prim, tris = bvh.download()
    assert np.array_equal(prim, prim_before)
    ref = orc.refit_bvh(orc.build_bvh4(mesh), moved)
    assert nodes.tobytes() == ref.nodes.tobytes()
    assert np.array_equal(tris.view(np.uint32), ref.triangles.view(np.uint32))


@pytest.mark.parametrize("name", ["single", "icosphere4", "torus", "soup", "displaced_256x128"])
def test_refit_bvh_matches_oracle(gwn, orc, name):
    mesh = MESHES[name]
    geom, bvh = build_gpu(gwn, mesh)
    moved = _deformed(mesh, 11)
    gwn.gwn_update_geometry(geom, moved.vx, moved.vy, moved.vz).raise_if_error()
    gwn.gwn_refit_bvh(geom, bvh).raise_if_error()
    ref = orc.refit_bvh(orc.build_bvh4(mesh), moved)
    info = bvh.info()
    nodes, prim, tris = bvh.download()
    assert nodes.tobytes() == ref.nodes.tobytes()
    assert np.array_equal(tris.view(np.uint32), ref.triangles.view(np.uint32))
    assert np.array_equal(info["root_bounds"].view(np.uint32), ref.root_bounds.view(np.uint32))
    assert info["internal_stack_bound"] == ref.internal_stack_bound
    helpers.check_bvh4(nodes, info["root_reference"], prim, tris, moved,
                       ref.internal_stack_bound, ref.packed_stack_bound)


def test_refit_bvh_then_moment_refit_equals_fresh_state(gwn, orc):
    """Refit to new positions, refit moments: bit-identical to the oracle's refit + moments, and
    a stale moment (formed before the refit) is rejected by the Taylor batch."""
    torch = _torch()
    mesh = MESHES["displaced_256x128"]
    geom, bvh = build_gpu(gwn, mesh)
    stale = gwn.MomentObject(1)
    gwn.gwn_refit_bvh_moment(bvh, stale).raise_if_error()
    moved = _deformed(mesh, 3)
    gwn.gwn_update_geometry(geom, moved.vx, moved.vy, moved.vz).raise_if_error()
    gwn.gwn_refit_bvh(geom, bvh).raise_if_error()
    q = [torch.zeros(8, device="cuda") for _ in range(3)]
    out = torch.empty(8, device="cuda")
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, stale, *q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
    mom = gwn.MomentObject(1)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    ref = orc.refit_bvh(orc.build_bvh4(mesh), moved)
    assert np.array_equal(mom.download(bvh.info()["node_count"]).view(np.uint32),
                          orc.refit_moment(ref, 1).view(np.uint32))


def test_refit_bvh_validation(gwn):
    mesh = MESHES["torus"]
    geom, bvh = build_gpu(gwn, mesh)
    other = gwn.GeometryObject()
    small = MESHES["two"]
    gwn.gwn_upload_geometry(other, small.vx, small.vy, small.vz, small.i0, small.i1, small.i2).raise_if_error()
    st = gwn.gwn_refit_bvh(other, bvh)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "triangle count differs" in st.message
    empty = gwn.BvhObject()
    st = gwn.gwn_refit_bvh(geom, empty)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "BVH accessor is invalid" in st.message
