"""Batched mode (gwn_batch_*): per-mesh results bit-identical to sequential single-mesh Taylor
batches with the same parameters (same trees, same accept/descend tests, same term order per
query, and terms contracted into FMAs only within each source expression, so every kernel rounds
them alike), in both traversals, before and after tuning, and across hipGraph replays with other
GPU work in between."""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu


def _affine(mesh, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    scale = rng.uniform(0.5, 2.0)
    shift = rng.uniform(-3, 3, size=3)
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float64) @ q.T * scale + shift
    return synth._mesh_from_arrays(v.astype(np.float32), np.stack([mesh.i0, mesh.i1, mesh.i2], 1))


def _meshes(count, seed=0):
    rng = np.random.default_rng(seed)
    out = [synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), [[0, 1, 2]])]
    for k in range(count - 1):
        tris = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        if k % 2 == 0:
            nu = max(3, int(np.sqrt(tris / 2)))
            base = synth.torus(nu, max(3, tris // (2 * nu)))
        else:
            nlon = max(4, int(np.sqrt(tris / 2)))
            base = synth.displaced_sphere(nlon, max(3, tris // (2 * nlon)), seed=k)
        out.append(_affine(base, 100 + k))
    return out


def _queries(mesh, n, seed):
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1)
    lo, hi = v.min(0), v.max(0)
    c, e = (lo + hi) / 2, (hi - lo) / 2 * 1.5 + 1e-3
    return synth.uniform_box_points(seed, n, c - e, c + e)


def _setup(gwn, dev, count, order=1, seed=0):
    import torch

    rng = np.random.default_rng(seed + 1)
    items, refs = [], []
    for k, mesh in enumerate(_meshes(count, seed)):
        geom = gwn.GeometryObject()
        gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
        bvh = gwn.BvhObject()
        method = gwn.BuildMethod.HPLOC if k % 3 else gwn.BuildMethod.LBVH
        gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method)).raise_if_error()
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        n = [0, 1, 63, 500, 70000][k] if k < 5 else int(np.exp(rng.uniform(np.log(64), np.log(20000))))
        q = [torch.from_numpy(a).to(dev) for a in _queries(mesh, n, 50 + k)]
        out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
        ref = torch.empty_like(out)
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ref, accuracy_scale=2.0).raise_if_error()
        items.append((bvh, mom, q[0], q[1], q[2], out))
        refs.append((geom, ref))
    torch.cuda.synchronize()
    return items, refs


def _check(items, refs):
    import torch

    torch.cuda.synchronize()
    for k, ((_, _, _, _, _, out), (_, ref)) in enumerate(zip(items, refs)):
        got, want = out.cpu().numpy(), ref.cpu().numpy()
        diff = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
        assert diff.size == 0, f"mesh {k}: {diff.size} results differ, max {np.abs(got - want).max():.3e}"


@pytest.mark.parametrize("order", [0, 1, 2])
def test_batch_matches_sequential_calls(gwn, dev, order):
    items, refs = _setup(gwn, dev, 24, order=order)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    batch.query().raise_if_error()
    _check(items, refs)
    t = batch.tuning()
    assert t["meshes"] == 24 and t["queries"] == sum(int(it[2].numel()) for it in items)


def test_batch_tuned_and_graph_replays(gwn, dev):
    import torch

    items, refs = _setup(gwn, dev, 40, order=1, seed=3)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    ms = batch.tune()
    assert ms > 0.0
    t = batch.tuning()
    assert t["block_size"] in (64, 128, 256) and t["chunk"] % t["block_size"] == 0 and t["traversal"] in (0, 1, 2)
    batch.query().raise_if_error()
    _check(items, refs)
    batch.capture(include_refit=True).raise_if_error()
    for rep in range(3):
        for it in items:
            it[5].fill_(float("nan"))  # eager work between replays
        torch.cuda.synchronize()
        batch.replay().raise_if_error()
        _check(items, refs)


def test_batch_other_tunings_and_validation(gwn, dev):
    items, refs = _setup(gwn, dev, 12, order=1, seed=5)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    # Traversal 2 picks per mesh by queries per triangle: 1e-3 sends every mesh with a query order
    # to the wave traversal, 1e6 none, and the middle values split this batch's meshes.
    shapes = ((64, 64, 1, 0, 1.0), (128, 512, 2, 0, 1.0), (256, 1024, 8, 0, 1.0),
              (64, 64, 1, 1, 1.0), (128, 256, 4, 1, 1.0), (256, 1024, 2, 1, 1.0),
              (256, 256, 4, 2, 1e-3), (128, 512, 2, 2, 0.5), (256, 1024, 8, 2, 2.0), (64, 64, 1, 2, 1e6))
    for block, chunk, bpc, trav, dens in shapes:
        batch.set_tuning(block, chunk, bpc, trav, wave_queries_per_triangle=dens).raise_if_error()
        t = batch.tuning()
        assert t["traversal"] == trav
        if trav == 2:
            assert t["wave_queries_per_triangle"] == pytest.approx(dens)
        for it in items:
            it[5].fill_(float("nan"))
        batch.query().raise_if_error()
        _check(items, refs)
    assert batch.set_tuning(100, 100, 1).error == gwn.Error.INVALID_ARGUMENT
    assert batch.set_tuning(64, 96, 1).error == gwn.Error.INVALID_ARGUMENT
    assert batch.set_tuning(64, 64, 1, 3).error == gwn.Error.INVALID_ARGUMENT
    assert batch.set_tuning(64, 64, 1, 2, wave_queries_per_triangle=0.0).error == gwn.Error.INVALID_ARGUMENT
    assert batch.set_tuning(64, 64, 1, 2, wave_queries_per_triangle=float("nan")).error == gwn.Error.INVALID_ARGUMENT
    assert batch.replay().error == gwn.Error.INVALID_ARGUMENT  # nothing captured
    mixed = list(items[:2])
    other = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(mixed[1][0], other).raise_if_error()
    mixed[1] = (mixed[1][0], other) + tuple(mixed[1][2:])
    with pytest.raises(Exception):
        gwn.BatchObject(mixed)


def test_batch_refuses_rebuilt_mesh_and_tracks_revisions(gwn, dev):
    """A replay after a mesh was rebuilt is refused (the captured pointers would be stale); a
    replay with refits leaves every moment usable with its BVH's current revision."""
    import torch

    items, refs = _setup(gwn, dev, 6, order=1, seed=9)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    batch.capture(include_refit=True).raise_if_error()
    geom = refs[2][0]
    gwn.gwn_refit_bvh(geom, items[2][0]).raise_if_error()  # bounds refit: new revision, same storage
    batch.replay().raise_if_error()
    torch.cuda.synchronize()
    _check(items, refs)
    out = torch.empty_like(items[2][5])
    gwn.gwn_compute_winding_number_taylor_batch(items[2][0], items[2][1], *items[2][2:5], out,
                                                accuracy_scale=2.0).raise_if_error()
    gwn.gwn_build_bvh(geom, items[2][0]).raise_if_error()  # rebuild: new storage
    st = batch.replay()
    assert st.error == gwn.Error.INVALID_ARGUMENT and "storage changed" in st.message
    assert batch.query().error == gwn.Error.INVALID_ARGUMENT


def test_batch_refuses_stale_moments_without_refit(gwn, dev):
    """After a bounds refit the moments no longer match their BVH's revision: an eager batch query
    and a replay of a graph captured without refits are refused; an eager batch refit, or a replay
    with refits, makes them usable again."""
    import torch

    items, refs = _setup(gwn, dev, 5, order=1, seed=11)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    batch.capture(include_refit=False).raise_if_error()
    batch.replay().raise_if_error()
    _check(items, refs)
    gwn.gwn_refit_bvh(refs[3][0], items[3][0]).raise_if_error()
    st = batch.query()
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
    st = batch.replay()
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
    batch.refit_moments().raise_if_error()
    batch.query().raise_if_error()
    batch.replay().raise_if_error()
    _check(items, refs)
    assert batch.status().is_ok()


def test_batch_capture_does_not_stamp_revisions(gwn, dev):
    """Capturing refits records them without running them: until a replay runs, the moments stay
    formed from the old bounds and a single-mesh query refuses them."""
    import torch

    items, refs = _setup(gwn, dev, 4, order=1, seed=13)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    bvh, mom = items[2][0], items[2][1]
    gwn.gwn_refit_bvh(refs[2][0], bvh).raise_if_error()
    batch.capture(include_refit=True).raise_if_error()
    out = torch.empty_like(items[2][5])
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *items[2][2:5], out, accuracy_scale=2.0)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
    batch.replay().raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *items[2][2:5], out, accuracy_scale=2.0).raise_if_error()
    torch.cuda.synchronize()
    _check(items, refs)


def test_batch_refuses_storage_reused_at_the_same_addresses(gwn, dev):
    """The pool may hand a rebuilt BVH or refit moment the addresses it had before; the batch still
    refuses it, because every swap of storage takes a new generation."""
    items, refs = _setup(gwn, dev, 3, order=1, seed=17)
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    geom = refs[1][0]
    for _ in range(2):
        gwn.gwn_build_bvh(geom, items[1][0]).raise_if_error()
    st = batch.query()
    assert st.error == gwn.Error.INVALID_ARGUMENT and "storage changed" in st.message
    items2, refs2 = _setup(gwn, dev, 3, order=2, seed=19)
    batch2 = gwn.BatchObject(items2, accuracy_scale=2.0)
    for _ in range(2):
        gwn.gwn_refit_bvh_moment(items2[2][0], items2[2][1]).raise_if_error()
    st = batch2.query()
    assert st.error == gwn.Error.INVALID_ARGUMENT and "storage changed" in st.message
    assert batch2.capture(include_refit=True).error == gwn.Error.INVALID_ARGUMENT


def test_batch_rejects_shared_moment_objects(gwn, dev):
    items, _ = _setup(gwn, dev, 3, order=1, seed=23)
    shared = list(items)
    shared[2] = (shared[1][0], shared[1][1]) + tuple(shared[2][2:])
    with pytest.raises(Exception, match="own moment"):
        gwn.BatchObject(shared, accuracy_scale=2.0)
