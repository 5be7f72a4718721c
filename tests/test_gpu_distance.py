"""GPU parity of the unsigned-distance batch (gwn_compute_unsigned_distance_batch).

Both sides run the same near-first culled traversal with unfused FP32 arithmetic and a
correctly rounded square root, so every distance must equal the oracle's replay bit for bit.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES, build_gpu

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _distance(gwn, dev, bvh, qx, qy, qz, band=float("inf"), sort=-1, stack=64, lanes=0, nearest=False):
    import torch

    out = torch.full((qx.shape[0],), -7.0, dtype=torch.float32, device=dev)
    near = torch.full((qx.shape[0],), 7, dtype=torch.int32, device=dev) if nearest else None
    gwn.gwn_compute_unsigned_distance_batch(
        bvh, _t(qx, dev), _t(qy, dev), _t(qz, dev), out, culling_band=band, nearest=near,
        config=gwn.QueryConfig(sort_queries=sort, stack_capacity=stack, lanes_per_point=lanes)).raise_if_error()
    torch.cuda.synchronize()
    if nearest:
        return out.cpu().numpy(), near.cpu().numpy().view(np.uint32)
    return out.cpu().numpy()


@pytest.mark.parametrize("lanes", [1, 4])
def test_distance_layered_kats(gwn, dev, lanes):
    """tests/unit_gwn_query_distance_ray.cu:222-270, 350-381 and 555-599."""
    f = np.float32
    _, one = build_gpu(gwn, synth.layered_triangles(1))
    assert _distance(gwn, dev, one, f([0.25]), f([0.25]), f([2]), lanes=lanes)[0] == 2.0
    assert _distance(gwn, dev, one, f([2]), f([0]), f([0]), lanes=lanes)[0] == 1.0
    assert _distance(gwn, dev, one, f([0.25]), f([0.25]), f([2]), band=0.5, lanes=lanes)[0] == 0.5
    assert _distance(gwn, dev, one, f([0.25]), f([0.25]), f([2]), band=-4.0, lanes=lanes)[0] == 0.0
    _, twelve = build_gpu(gwn, synth.layered_triangles(12))
    d = _distance(gwn, dev, twelve, f([0.25, 0.25, 0.0]), f([0.25, 0.25, 0.25]), f([10.25, 30, 10.25]), band=2.0,
                  lanes=lanes)
    assert d.tolist() == [0.25, 2.0, 0.25]
    _, four = build_gpu(gwn, synth.layered_triangles(4))
    info = four.info()
    assert info["internal_stack_bound"] == 0 and info["packed_stack_bound"] == 3
    assert _distance(gwn, dev, four, f([0.25]), f([0.25]), f([3.5]), stack=1, lanes=lanes)[0] == 0.5


@pytest.mark.parametrize("lanes", [1, 4], ids=["lane", "quad"])
@pytest.mark.parametrize("sort", [0, 1], ids=["input_order", "morton_order"])
@pytest.mark.parametrize("name", ["torus", "soup", "displaced_256x128", "icosphere4"])
def test_distance_matches_oracle_bitwise(gwn, orc, dev, name, sort, lanes):
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    ref_bvh = orc.build_bvh4(mesh)
    qx, qy, qz = synth.mixed_queries(mesh, 8192, seed=41, band=0.01) if name != "soup" else \
        synth.uniform_box_points(41, 8192, (-3.0,) * 3, (3.0,) * 3)
    for band in (float("inf"), 0.05):
        got = _distance(gwn, dev, bvh, qx, qy, qz, band=band, sort=sort, lanes=lanes)
        ref = orc.unsigned_distance_f32(ref_bvh, qx, qy, qz, culling_band=band)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"band {band}"


@pytest.mark.parametrize("name", ["torus", "soup", "displaced_256x128", "icosphere4"])
def test_distance_wave_traversal_within_ulps(gwn, orc, dev, name):
    """sort_queries = 2 (wave-cooperative traversal): every distance is within 4 FP32 ulps of the
    reference-order replay, NaN and zero patterns equal, and at least 90% identical.  Only
    pruning near ties of box and triangle distances depends on the visiting order."""
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    ref_bvh = orc.build_bvh4(mesh)
    qx, qy, qz = synth.mixed_queries(mesh, 8192, seed=43, band=0.01) if name != "soup" else \
        synth.uniform_box_points(43, 8192, (-3.0,) * 3, (3.0,) * 3)
    for band in (float("inf"), 0.05):
        got = _distance(gwn, dev, bvh, qx, qy, qz, band=band, sort=2).astype(np.float64)
        ref = orc.unsigned_distance_f32(ref_bvh, qx, qy, qz, culling_band=band).astype(np.float64)
        assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.array_equal(got == 0, ref == 0)
        ok = ~np.isnan(ref)
        ulp = np.spacing(ref[ok].astype(np.float32)).astype(np.float64)
        assert np.all(np.abs(got[ok] - ref[ok]) <= 4 * ulp), f"band {band}"
        assert np.mean(got[ok] == ref[ok]) >= 0.9


def test_distance_on_surface_points_is_zero_or_tiny(gwn, orc, dev):
    mesh = MESHES["displaced_256x128"]
    _, bvh = build_gpu(gwn, mesh)
    p, _ = synth.surface_points(mesh, 3, 4096)
    qx, qy, qz = (np.ascontiguousarray(p[:, k].astype(np.float32)) for k in range(3))
    got = _distance(gwn, dev, bvh, qx, qy, qz)
    assert np.all(got >= 0) and got.max() < 1e-5
    ref = orc.unsigned_distance_f32(orc.build_bvh4(mesh), qx, qy, qz)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_distance_large_mesh_subsample(gwn, orc, dev):
    """1,048,576-triangle mesh, 262,144 queries through the Morton-ordered path; a strided
    subsample checked bit for bit against the oracle, and the one- and four-lane traversals equal
    at every point (distances and nearest records)."""
    mesh = synth.config_large_mesh()
    _, bvh = build_gpu(gwn, mesh)
    qx, qy, qz = synth.mixed_queries(mesh, 262_144, seed=5, band=0.01)
    got, rec = _distance(gwn, dev, bvh, qx, qy, qz, lanes=4, nearest=True)
    idx = np.arange(0, 262_144, 64)
    ref = orc.unsigned_distance_f32(orc.build_bvh4(mesh), qx[idx], qy[idx], qz[idx])
    assert np.array_equal(got[idx].view(np.uint32), ref.view(np.uint32))
    assert np.all(np.isfinite(got))
    lane, lane_rec = _distance(gwn, dev, bvh, qx, qy, qz, lanes=1, nearest=True)
    assert np.array_equal(got.view(np.uint32), lane.view(np.uint32))
    assert np.array_equal(rec, lane_rec)


@pytest.mark.parametrize("name", ["torus", "soup", "displaced_256x128"])
def test_distance_quad_records_equal_lane_records(gwn, dev, name):
    """Four lanes per point against one lane per point, in input and Morton order, with and
    without a band: identical distances and nearest records (first triangle attaining the minimum
    in the traversal's order), including zero, band and overflow codes."""
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    qx, qy, qz = synth.mixed_queries(mesh, 20_000, seed=47, band=0.01) if name != "soup" else \
        synth.uniform_box_points(47, 20_000, (-3.0,) * 3, (3.0,) * 3)
    if name == "torus":  # points on vertices: distance zero
        qx[:64], qy[:64], qz[:64] = mesh.vx[:64], mesh.vy[:64], mesh.vz[:64]
    for sort in (0, 1):
        for band in (float("inf"), 0.05, 0.0):
            d1, r1 = _distance(gwn, dev, bvh, qx, qy, qz, band=band, sort=sort, lanes=1, nearest=True)
            d4, r4 = _distance(gwn, dev, bvh, qx, qy, qz, band=band, sort=sort, lanes=4, nearest=True)
            assert np.array_equal(d1.view(np.uint32), d4.view(np.uint32)), (sort, band)
            assert np.array_equal(r1, r4), (sort, band)
    info = bvh.info()
    small = int(info["internal_stack_bound"]) - 1
    if small >= 1:  # a stack below the bound is rejected up front: no overflow path to compare
        st = gwn.gwn_compute_unsigned_distance_batch(
            bvh, *[_t(a, dev) for a in (qx, qy, qz)], _t(np.zeros_like(qx), dev),
            config=gwn.QueryConfig(stack_capacity=small, lanes_per_point=4))
        assert st.error == gwn.Error.INVALID_ARGUMENT


def test_distance_validation_contract(gwn, dev):
    """gwn_query.cuh:150-168."""
    import torch

    _, bvh = build_gpu(gwn, MESHES["torus"])
    q = [torch.zeros(4, device=dev) for _ in range(3)]
    out = torch.zeros(4, device=dev)
    st = gwn.gwn_compute_unsigned_distance_batch(gwn.BvhObject(), *q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "no queryable data" in st.message
    st = gwn.gwn_compute_unsigned_distance_batch(bvh, *q, torch.zeros(3, device=dev))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Output span size" in st.message
    st = gwn.gwn_compute_unsigned_distance_batch(bvh, *q, out, config=gwn.QueryConfig(stack_capacity=1))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Traversal stack capacity" in st.message
    st = gwn.gwn_compute_unsigned_distance_batch(bvh, *q, out, config=gwn.QueryConfig(lanes_per_point=2))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Lanes per point" in st.message
    empty = torch.zeros(0, device=dev)
    assert gwn.gwn_compute_unsigned_distance_batch(bvh, empty, empty, empty, empty).is_ok()
