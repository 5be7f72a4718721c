"""Batched mode over geometries (gwn_batch_create_built): every mesh's tree, records and moments
come from one segmented pass over all meshes, and must equal gwn_build_bvh (the same builder --
H-PLOC with the same radius, or LBVH -- and width) + gwn_refit_bvh_moment of that mesh alone bit
for bit; the query results must equal the
single-mesh Taylor batch's bit for bit, eagerly and in graph replays that rebuild every tree from
updated vertices.  The meshes cover one and two triangles, repeated triangles (equal Morton codes,
so ties break by mesh-local position), uneven sizes, and enough meshes for a two-pass mesh sort."""

import numpy as np
import pytest

from smallgwn_amd import synth

from .test_gpu_batch import _affine, _queries

pytestmark = pytest.mark.gpu


def _meshes(count, seed=0, max_tris=20000):
    rng = np.random.default_rng(seed)
    out = [
        synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), [[0, 1, 2]]),
        synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]]), [[0, 1, 2], [1, 3, 2]]),
    ]
    # The same triangles four times over: runs of equal codes.
    s = synth.displaced_sphere(12, 6, seed=5)
    f = np.stack([s.i0, s.i1, s.i2], 1)
    out.append(synth._mesh_from_arrays(np.stack([s.vx, s.vy, s.vz], 1), np.concatenate([f] * 4)))
    for k in range(count - len(out)):
        tris = int(np.exp(rng.uniform(np.log(16), np.log(max_tris))))
        if k % 2 == 0:
            nu = max(3, int(np.sqrt(tris / 2)))
            base = synth.torus(nu, max(3, tris // (2 * nu)))
        else:
            nlon = max(4, int(np.sqrt(tris / 2)))
            base = synth.displaced_sphere(nlon, max(3, tris // (2 * nlon)), seed=k)
        out.append(_affine(base, 200 + k))
    return out[:count]


def _setup(gwn, dev, meshes, seed=0):
    import torch

    rng = np.random.default_rng(seed + 1)
    geoms, items = [], []
    for k, mesh in enumerate(meshes):
        geom = gwn.GeometryObject()
        gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
        n = [0, 70, 1000][k] if k < 3 else int(np.exp(rng.uniform(np.log(1), np.log(30000))))
        q = [torch.from_numpy(a).to(dev) for a in _queries(mesh, n, 70 + k)]
        out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
        geoms.append(geom)
        items.append((geom, q[0], q[1], q[2], out))
    return geoms, items


def _sequential(gwn, geom, q, width, radius, order, method=None):
    import torch

    method = gwn.BuildMethod.HPLOC if method is None else method
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method, hploc_search_radius=radius,
                                                  width=width)).raise_if_error()
    mom = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    ref = torch.empty_like(q[0])
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ref, accuracy_scale=2.0).raise_if_error()
    return bvh, mom, ref


def _same_tree(k, got, want):
    gb, gm = got
    wb, wm = want
    gi, wi = gb.info(), wb.info()
    for key in ("node_count", "triangle_count", "internal_stack_bound", "packed_stack_bound", "max_depth", "width",
                "root_reference"):
        assert gi[key] == wi[key], f"mesh {k}: {key} {gi[key]} != {wi[key]}"
    assert np.array_equal(gi["root_bounds"].view(np.uint32), wi["root_bounds"].view(np.uint32)), f"mesh {k}: root"
    for a, b, what in zip(gb.download(), wb.download(), ("nodes", "primitive order", "records")):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), f"mesh {k}: {what}"
    n = wi["node_count"]
    assert np.array_equal(gm.download(n).view(np.uint32), wm.download(n).view(np.uint32)), f"mesh {k}: moments"


def _same_results(k, got, want):
    g, w = got.cpu().numpy(), want.cpu().numpy()
    diff = np.flatnonzero(g.view(np.uint32) != w.view(np.uint32))
    assert diff.size == 0, f"mesh {k}: {diff.size} of {g.size} results differ"


@pytest.mark.parametrize("lbvh,count,width,radius,order", [
    (False, 40, 4, 8, 1), (False, 24, 2, 1, 0), (False, 24, 3, 4, 2), (False, 300, 4, 8, 1),
    (True, 40, 4, 8, 1), (True, 24, 2, 8, 2), (True, 300, 3, 8, 0)])
def test_built_batch_equals_single_mesh_builds(gwn, dev, lbvh, count, width, radius, order):
    import torch

    method = gwn.BuildMethod.LBVH if lbvh else gwn.BuildMethod.HPLOC
    meshes = _meshes(count, seed=count + width, max_tris=20000 if count < 100 else 4000)
    geoms, items = _setup(gwn, dev, meshes, seed=count)
    batch = gwn.BuiltBatchObject(items, width=width, hploc_radius=radius, order=order, accuracy_scale=2.0,
                                 method=method)
    cfg = batch.build_config()
    assert (cfg["method"], cfg["width"], cfg["hploc_search_radius"], cfg["order"]) == (method, width, radius, order)
    assert 1 <= cfg["level_launches"] <= 128
    batch.query().raise_if_error()
    torch.cuda.synchronize()
    for k, (geom, qx, qy, qz, out) in enumerate(items):
        want = _sequential(gwn, geom, (qx, qy, qz), width, radius, order, method)
        _same_tree(k, batch.export(k), want[:2])
        _same_results(k, out, want[2])
    batch.status().raise_if_error()


@pytest.mark.parametrize("lbvh", [False, True])
def test_built_batch_graph_rebuilds_updated_geometry(gwn, dev, lbvh):
    import torch

    method = gwn.BuildMethod.LBVH if lbvh else gwn.BuildMethod.HPLOC
    meshes = _meshes(32, seed=9)
    geoms, items = _setup(gwn, dev, meshes, seed=9)
    batch = gwn.BuiltBatchObject(items, width=4, hploc_radius=8, order=1, accuracy_scale=2.0, method=method)
    batch.capture(include_refit=True).raise_if_error()
    rng = np.random.default_rng(5)
    for rep in range(3):
        # Move every vertex a little (the topology of the trees changes), in place.
        for geom, mesh in zip(geoms, meshes):
            d = rng.normal(scale=0.02, size=(3, mesh.vx.size)).astype(np.float32)
            gwn.gwn_update_geometry(geom, mesh.vx + d[0], mesh.vy + d[1], mesh.vz + d[2]).raise_if_error()
        for it in items:
            it[4].fill_(float("nan"))
        torch.cuda.synchronize()
        batch.replay().raise_if_error()
        batch.status().raise_if_error()
        for k, (geom, qx, qy, qz, out) in enumerate(items):
            want = _sequential(gwn, geom, (qx, qy, qz), 4, 8, 1, method)
            _same_results(k, out, want[2])
            if k % 7 == 0:
                _same_tree(k, batch.export(k), want[:2])
    # An eager pass gives the same as the replay.
    for it in items:
        it[4].fill_(float("nan"))
    batch.build().raise_if_error()
    batch.query().raise_if_error()
    torch.cuda.synchronize()
    for k, (geom, qx, qy, qz, out) in enumerate(items):
        _same_results(k, out, _sequential(gwn, geom, (qx, qy, qz), 4, 8, 1, method)[2])


def test_built_batch_tunes_build_and_launch(gwn, dev):
    import torch

    meshes = _meshes(20, seed=4, max_tris=8000)
    geoms, items = _setup(gwn, dev, meshes, seed=4)
    batch = gwn.BuiltBatchObject(items, width=4, hploc_radius=8, order=1, accuracy_scale=2.0)
    assert batch.tune() > 0.0
    assert batch.tune_build() > 0.0
    cfg = batch.build_config()
    assert cfg["width"] in (2, 3, 4) and cfg["hploc_search_radius"] in (1, 2, 4, 8)
    assert cfg["method"] in (gwn.BuildMethod.HPLOC, gwn.BuildMethod.LBVH)
    t = batch.tuning()
    assert t["stack_capacity"] in (32, 64, 128)
    batch.capture(include_refit=True).raise_if_error()
    batch.replay().raise_if_error()
    batch.status().raise_if_error()
    torch.cuda.synchronize()
    for k, (geom, qx, qy, qz, out) in enumerate(items):
        want = _sequential(gwn, geom, (qx, qy, qz), cfg["width"], cfg["hploc_search_radius"], 1, cfg["method"])
        assert t["stack_capacity"] >= want[0].info()["internal_stack_bound"]
        _same_results(k, out, want[2])


def test_built_batch_validation(gwn, dev):
    meshes = _meshes(6, seed=2, max_tris=2000)
    geoms, items = _setup(gwn, dev, meshes, seed=2)
    batch = gwn.BuiltBatchObject(items, width=4, hploc_radius=8, order=1)
    st = batch.refit_moments()
    assert not st.is_ok() and "gwn_batch_build" in st.message
    assert not batch.set_build(4, 9).is_ok()
    assert not batch.set_build(5, 8).is_ok()
    assert not batch.set_tuning(256, 256, 4, 1, stack_capacity=48).is_ok()
    with pytest.raises(Exception):
        batch.export(len(items))
    empty = gwn.GeometryObject()
    with pytest.raises(Exception):
        gwn.BuiltBatchObject([(empty,) + items[0][1:]])
    # Re-uploading a mesh replaces its storage: the batch refuses to run.
    m = meshes[3]
    gwn.gwn_upload_geometry(geoms[3], m.vx, m.vy, m.vz, m.i0, m.i1, m.i2).raise_if_error()
    st = batch.build()
    assert not st.is_ok() and "geometry storage changed" in st.message
    assert not batch.query().is_ok()


def _soup(centers, e):
    """One small triangle (c, c + e x, c + e y) per center, as a triangle soup."""
    c = np.asarray(centers, dtype=np.float64)
    v = np.stack([c, c + [e, 0.0, 0.0], c + [0.0, e, 0.0]], 1).reshape(-1, 3).astype(np.float32)
    return synth._mesh_from_arrays(v, np.arange(v.shape[0], dtype=np.int64).reshape(-1, 3))


@pytest.mark.parametrize("lbvh", [False, True])
def test_built_batch_flags_a_deeper_rebuild_and_replans(gwn, dev, lbvh):
    """A geometry update whose tree is deeper than the levels planned at create: the pass flags it
    (status), set_build re-plans, and the next pass equals the single-mesh build again."""
    import torch

    method = gwn.BuildMethod.LBVH if lbvh else gwn.BuildMethod.HPLOC
    grid = [(i / 4.0, j / 12.0, 0.5) for i in range(5) for j in range(13)]  # 65 triangles
    meshes = [_soup(grid, 0.01), _soup(grid[::-1], 0.02)]
    # The same 65 triangles moved to single-axis points 2^-k (k = 1..21 on each axis) plus the
    # corners (0,0,0) and (1,1,1): every point's Morton code has its own leading bit, so the
    # binary tree is a chain about 64 deep, against about 8 for the grid.
    chain = [(0.0, 0.0, 0.0), (1.0, 1.0, 1.0)]
    for axis in range(3):
        for k in range(1, 22):
            p = [0.0, 0.0, 0.0]
            p[axis] = 2.0 ** -k
            chain.append(tuple(p))
    deep = _soup(chain, 1e-9)
    items = []
    for k, mesh in enumerate(meshes):
        geom = gwn.GeometryObject()
        gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
        q = [torch.from_numpy(a).to(dev) for a in _queries(mesh, 300, 90 + k)]
        items.append((geom, q[0], q[1], q[2], torch.full((300,), float("nan"), dtype=torch.float32, device=dev)))
    batch = gwn.BuiltBatchObject(items, width=4, hploc_radius=8, order=1, accuracy_scale=2.0, method=method)
    planned = batch.build_config()["level_launches"]
    gwn.gwn_update_geometry(items[0][0], deep.vx, deep.vy, deep.vz).raise_if_error()
    batch.build().raise_if_error()
    st = batch.status()
    if lbvh:  # the binary chain is far deeper than the binary-depth launches planned for the grid
        assert not st.is_ok() and "re-plan" in st.message, st.message
    elif not st.is_ok():
        assert "re-plan" in st.message or "stack" in st.message, st.message
    batch.set_build(4, 8, method).raise_if_error()
    assert batch.build_config()["level_launches"] >= planned
    want0 = _sequential(gwn, items[0][0], items[0][1:4], 4, 8, 1, method)
    for it in items:
        it[4].fill_(float("nan"))
    batch.build().raise_if_error()
    batch.query().raise_if_error()
    batch.status().raise_if_error()
    torch.cuda.synchronize()
    for k, (geom, qx, qy, qz, out) in enumerate(items):
        want = want0 if k == 0 else _sequential(gwn, geom, (qx, qy, qz), 4, 8, 1, method)
        _same_tree(k, batch.export(k), want[:2])
        _same_results(k, out, want[2])
