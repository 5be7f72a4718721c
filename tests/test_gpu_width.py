"""Trees of width 2 and 3 (gwn_bvh_object<Width, float, uint32_t>) through the C ABI.

The reference's builder suite runs widths 2, 3 and 4 (tests/integration_gwn_builder.cu:256-298).
Here a narrower tree is the same sort and binary hierarchy collapsed with at most Width children
per node (detail/gwn_bvh_build_collapse.cuh:163-296), stored four-wide with the extra slots empty.
Checked against the oracle's collapse of the same width, bit for bit: nodes in the reference's
Width-wide layout, primitive order, records, stack bounds, bounds refit and moments of orders
0-2; and the queries (exact, Taylor and gradient with the oracle's replay counters, unsigned
distance, first hits, antipodal winding) on those trees equal the oracle's replays of the same
traversals.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests import mesh_e2e

pytestmark = pytest.mark.gpu

MESHES = {
    "torus": lambda: synth.torus(40, 20),
    "displaced": lambda: synth.displaced_sphere(96, 48),
    "soup": lambda: synth.open_soup(48, 24),
}


def _narrow(nodes4, width):
    return np.ascontiguousarray(nodes4["children"][:, :width]).tobytes()


def _build(gwn, mesh, method, width, radius=8):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method, hploc_search_radius=radius, width=width)).raise_if_error()
    return geom, bvh


@pytest.mark.parametrize("width", [2, 3])
@pytest.mark.parametrize("method", ["lbvh", "hploc"])
@pytest.mark.parametrize("name", list(MESHES))
def test_narrow_build_refit_and_moments_match_oracle(gwn, orc, name, method, width):
    print(f"[width] {name} {method} W={width}", flush=True)
    mesh = MESHES[name]()
    bm = gwn.BuildMethod.LBVH if method == "lbvh" else gwn.BuildMethod.HPLOC
    geom, bvh = _build(gwn, mesh, bm, width)
    ref = orc.build_bvh4(mesh, method, 8, width=width)
    info = bvh.info()
    assert info["width"] == width and info["node_count"] == ref.node_count
    assert (info["internal_stack_bound"], info["packed_stack_bound"]) == (ref.internal_stack_bound,
                                                                         ref.packed_stack_bound)
    nodes, prim, tris = bvh.download()
    assert nodes.tobytes() == _narrow(ref.nodes, width)
    assert np.array_equal(prim, ref.primitive_indices)
    assert np.array_equal(tris.view(np.uint32), ref.triangles.view(np.uint32))
    mesh_e2e.check_internal_nodes_complete(bvh, mesh.triangle_count)
    for order in (0, 1, 2):
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        got = mom.download(info["node_count"])
        want = orc.refit_moment(ref, order)
        assert got.shape[2] == width
        assert np.array_equal(got.view(np.uint32), np.ascontiguousarray(want[:, :, :width]).view(np.uint32))
        assert not want[:, :, width:].any()
    # Bounds refit after a deformation.
    rng = np.random.default_rng(5)
    moved = synth._mesh_from_arrays(
        np.stack([mesh.vx, mesh.vy, mesh.vz], 1) + rng.normal(scale=0.01, size=(mesh.vertex_count, 3)).astype(np.float32),
        np.stack([mesh.i0, mesh.i1, mesh.i2], 1))
    gwn.gwn_update_geometry(geom, moved.vx, moved.vy, moved.vz).raise_if_error()
    gwn.gwn_refit_bvh(geom, bvh).raise_if_error()
    rref = orc.refit_bvh(ref, moved)
    nodes, _, tris = bvh.download()
    assert nodes.tobytes() == _narrow(rref.nodes, width)
    assert np.array_equal(tris.view(np.uint32), rref.triangles.view(np.uint32))


@pytest.mark.parametrize("width", [2, 3])
def test_queries_on_narrow_trees_match_oracle_replays(gwn, orc, dev, width):
    import torch

    print(f"[width] queries W={width}", flush=True)
    mesh = synth.open_soup(48, 24)
    geom, bvh = _build(gwn, mesh, gwn.BuildMethod.HPLOC, width)
    ref = orc.build_bvh4(mesh, "hploc", 8, width=width)
    stack = max(64, bvh.info()["internal_stack_bound"], bvh.info()["packed_stack_bound"])
    cfg = gwn.QueryConfig(stack_capacity=stack)
    n = 4096
    q = synth.uniform_box_points(77, n, (-1.5,) * 3, (1.5,) * 3)
    d = [torch.from_numpy(a).to(dev) for a in q]
    for order in (1, 2):
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        rmom = orc.refit_moment(ref, order)
        for sort in (0, 1):
            out = torch.empty(n, device=dev)
            cnt = torch.zeros(3 * n, dtype=torch.int32, device=dev)
            gwn.gwn_compute_winding_number_taylor_batch(
                bvh, mom, *d, out, accuracy_scale=2.0, counters=cnt,
                config=gwn.QueryConfig(stack_capacity=stack, sort_queries=sort)).raise_if_error()
            torch.cuda.synchronize()
            want, wcnt = orc.winding_taylor_f32(ref, rmom, order, *q, accuracy_scale=2.0, counters=True,
                                                stack_capacity=stack)
            assert np.array_equal(cnt.cpu().numpy().reshape(-1, 3).astype(np.uint64), wcnt)
            got = out.cpu().numpy().astype(np.float64)
            assert (np.abs(got - want) / (1 + np.abs(want))).max() <= 1e-5
        g = [torch.empty(n, device=dev) for _ in range(3)]
        gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *d, *g, accuracy_scale=2.0, config=cfg).raise_if_error()
        torch.cuda.synchronize()
        _, *gw = orc.winding_gradient_taylor_f32(ref, rmom, order, *q, accuracy_scale=2.0, stack_capacity=stack)
        gw = np.stack(gw, 1).astype(np.float64)
        gg = np.stack([t.cpu().numpy() for t in g], 1).astype(np.float64)
        scale = 1 + np.abs(gw).max(axis=1, keepdims=True)
        assert (np.abs(gg - gw) / scale).max() <= 1e-4
    dist = torch.empty(n, device=dev)
    for lanes in (1, 4):
        gwn.gwn_compute_unsigned_distance_batch(
            bvh, *d, dist, config=gwn.QueryConfig(stack_capacity=stack, lanes_per_point=lanes)).raise_if_error()
        torch.cuda.synchronize()
        assert np.array_equal(dist.cpu().numpy().view(np.uint32),
                              orc.unsigned_distance_f32(ref, *q, stack_capacity=stack).view(np.uint32))
    rays = mesh_e2e.fixture_rays(mesh)
    hits = torch.empty(8 * rays[0].shape[0], dtype=torch.int32, device=dev)
    gwn.gwn_compute_ray_first_hit_batch(bvh, *[torch.from_numpy(a).to(dev) for a in rays], hits,
                                        config=cfg).raise_if_error()
    torch.cuda.synchronize()
    assert hits.cpu().numpy().tobytes() == orc.ray_first_hit_f32(ref, *rays, stack_capacity=stack).tobytes()
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain).raise_if_error()
    anti = torch.empty(n, device=dev)
    gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *d, anti, config=cfg).raise_if_error()
    torch.cuda.synchronize()
    want = orc.winding_antipodal_f32(ref, mesh, orc.boundary_chain(mesh), *q, stack_capacity=stack)
    got = anti.cpu().numpy()
    assert np.array_equal(np.isfinite(got), np.isfinite(want))
    ok = np.isfinite(want)
    assert np.abs(got[ok].astype(np.float64) - want[ok]).max() <= 1e-5


@pytest.mark.parametrize("width", [2, 3, 4])
def test_builder_contract_on_fixtures_at_every_width(gwn, orc, dev, width):
    """tests/integration_gwn_builder.cu:256-298: every builder and width on the fixture meshes and
    the equal-bounds counts, exact and descended Taylor against the FP64 exact winding."""
    import os

    from smallgwn_amd import ply

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")
    for path in ply.collect_ply_model_paths(here):
        mesh_e2e.builder_contract(gwn, orc, dev, ply.load_ply_mesh(path), width=width)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    q = (np.array([0.2, 0.7, -0.3], np.float32), np.array([0.2, 0.1, 0.3], np.float32),
         np.array([1.0, -2.0, 0.5], np.float32))
    for count in (15, 16, 17, 33, 65, 129):
        mesh = synth._mesh_from_arrays(v, np.tile([[0, 1, 2]], (count, 1)))
        mesh_e2e.builder_contract(gwn, orc, dev, mesh, methods=(gwn.BuildMethod.HPLOC,), radii=(1, 4, 8),
                                  queries=q, width=width)
