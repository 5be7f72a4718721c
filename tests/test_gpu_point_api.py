"""Device point API (include/gwn/gwn_device.hpp) in user kernels, against the batches.

tests/cpp/gwn_point_api.hip evaluates one query per thread through the reference-named device
functions from the objects' typed accessors.  Every result must equal the corresponding batch's bit
for bit (the per-lane traversals: sort_queries = 0, one lane per point), since both run the same
per-query code.  With a one-entry stack the traversals overflow: the callback fires for exactly the
queries whose result is NaN (distance, Taylor, antipodal) or whose hit has status k_overflow, and
the other queries still equal the batch (tests/unit_gwn_query_distance_ray.cu:86-137,
tests/unit_gwn_query_antipodal.cu:494-547).
"""

import ctypes
import os

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libgwn_point_api.so")


@pytest.fixture(scope="module")
def point(gwn):
    if not os.path.exists(SO):
        pytest.fail(f"{SO} is not built (tests/cpp/build.py)")
    lib = ctypes.CDLL(SO)
    lib.gwn_point_run.restype = ctypes.c_int
    vp, f32 = ctypes.c_void_p, ctypes.c_float
    lib.gwn_point_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  ctypes.c_uint64, f32, f32, vp, vp, vp, vp, vp, vp]
    return lib


def _h(obj):
    return obj.handle if obj is not None else None


def _run(gwn, lib, dev, op, q, order=0, cap=64, geom=None, bvh=None, mom=None, chain=None, a=0.0, b=0.0, rays=False):
    import torch

    n = q[0].numel()
    out = torch.full((8 * n,) if rays else (n,), -7, dtype=torch.int32 if rays else torch.float32, device=dev)
    g = [torch.full((n,), -7.0, device=dev) for _ in range(3)]
    flags = torch.zeros(n, dtype=torch.int32, device=dev)
    ptr = [t.data_ptr() for t in q] + [None] * (6 - len(q))
    rc = lib.gwn_point_run(op, order, cap, _h(geom), _h(bvh), _h(mom), _h(chain), *ptr, n, a, b, out.data_ptr(),
                           g[0].data_ptr(), g[1].data_ptr(), g[2].data_ptr(), flags.data_ptr(),
                           torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    res = out.cpu().numpy()
    return (res.view(gwn.RAY_HIT_DTYPE) if rays else res), np.stack([t.cpu().numpy() for t in g]), flags.cpu().numpy()


def _objects(gwn, mesh, order=None, chain=False):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh).raise_if_error()
    mom = None
    if order is not None:
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    c = None
    if chain:
        c = gwn.BoundaryChainObject()
        gwn.gwn_build_boundary_chain(geom, c).raise_if_error()
    return geom, bvh, mom, c


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _queries(dev, n=4096, seed=3):
    import torch

    return [torch.from_numpy(a).to(dev) for a in synth.uniform_box_points(seed, n, (-1.5,) * 3, (1.5,) * 3)]


def test_point_exact_and_taylor_equal_batches(gwn, point, dev):
    import torch

    mesh = synth.displaced_sphere(64, 32)
    q = _queries(dev)
    n = q[0].numel()
    geom, bvh, _, _ = _objects(gwn, mesh)
    got, _, _ = _run(gwn, point, dev, 0, q, bvh=bvh)
    want = torch.empty(n, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(bvh, *q, want).raise_if_error()
    torch.cuda.synchronize()
    assert np.array_equal(_bits(got), _bits(want.cpu().numpy()))
    cfg = gwn.QueryConfig(sort_queries=0)
    for order in (0, 1, 2):
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        w, g, flags = _run(gwn, point, dev, 1, q, order=order, bvh=bvh, mom=mom, a=2.0)
        ww = torch.empty(n, device=dev)
        gg = [torch.empty(n, device=dev) for _ in range(3)]
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ww, accuracy_scale=2.0, config=cfg).raise_if_error()
        gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *q, *gg, accuracy_scale=2.0, config=cfg).raise_if_error()
        torch.cuda.synchronize()
        assert not flags.any()
        assert np.array_equal(_bits(w), _bits(ww.cpu().numpy())), order
        assert np.array_equal(_bits(g), _bits(np.stack([t.cpu().numpy() for t in gg]))), order


def test_point_distance_and_rays_equal_batches(gwn, orc, point, dev):
    import torch

    from tests import mesh_e2e

    mesh = synth.torus(48, 24)
    q = _queries(dev, seed=4)
    n = q[0].numel()
    geom, bvh, _, _ = _objects(gwn, mesh)
    for band in (float("inf"), 0.1):
        got, _, flags = _run(gwn, point, dev, 2, q, bvh=bvh, a=band)
        want = torch.empty(n, device=dev)
        gwn.gwn_compute_unsigned_distance_batch(bvh, *q, want, culling_band=band,
                                                config=gwn.QueryConfig(sort_queries=0, lanes_per_point=1)).raise_if_error()
        torch.cuda.synchronize()
        assert not flags.any()
        assert np.array_equal(_bits(got), _bits(want.cpu().numpy())), band
    rays = [torch.from_numpy(a).to(dev) for a in mesh_e2e.fixture_rays(mesh)]
    m = rays[0].numel()
    got, _, flags = _run(gwn, point, dev, 3, rays, bvh=bvh, a=0.0, b=float("inf"), rays=True)
    want = torch.empty(8 * m, dtype=torch.int32, device=dev)
    gwn.gwn_compute_ray_first_hit_batch(bvh, *rays, want).raise_if_error()
    torch.cuda.synchronize()
    assert not flags.any()
    assert got.tobytes() == want.cpu().numpy().tobytes()
    assert (got["status"] == gwn.RayStatus.HIT).any() and (got["status"] == gwn.RayStatus.MISS).any()


def test_point_antipodal_equals_batch(gwn, point, dev):
    import torch

    mesh = synth.open_soup(32, 16)
    q = _queries(dev, n=2048, seed=5)
    n = q[0].numel()
    geom, bvh, _, chain = _objects(gwn, mesh, chain=True)
    w, g, flags = _run(gwn, point, dev, 4, q, geom=geom, bvh=bvh, chain=chain)
    ww = torch.empty(n, device=dev)
    gg = [torch.empty(n, device=dev) for _ in range(3)]
    gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *q, ww).raise_if_error()
    gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *q, *gg).raise_if_error()
    torch.cuda.synchronize()
    assert not flags.any()
    assert np.array_equal(_bits(w), _bits(ww.cpu().numpy()))
    assert np.array_equal(_bits(g), _bits(np.stack([t.cpu().numpy() for t in gg])))


def test_overflow_callback_contract(gwn, point, dev):
    """A one-entry stack: the callback fires exactly where the result reports the overflow."""
    import torch

    mesh = synth.displaced_sphere(64, 32)
    # Near queries descend into two internal children somewhere and overflow a one-entry stack;
    # far ones are accepted at the root's children and never push.
    near = synth.uniform_box_points(6, 2048, (-1.5,) * 3, (1.5,) * 3)
    far = synth.uniform_box_points(7, 2048, (-20.0,) * 3, (20.0,) * 3)
    q = [torch.from_numpy(np.concatenate([a, b])).to(dev) for a, b in zip(near, far)]
    n = q[0].numel()
    geom, bvh, mom, _ = _objects(gwn, mesh, order=1)
    assert bvh.info()["internal_stack_bound"] > 1
    w, g, flags = _run(gwn, point, dev, 1, q, order=1, cap=1, bvh=bvh, mom=mom, a=2.0)
    over = flags.astype(bool)
    assert over.any() and (~over).any()
    assert np.array_equal(np.isnan(w), over)
    assert np.isnan(g[:, over]).all() and np.isfinite(g[:, ~over]).all()
    ww = torch.empty(n, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ww, accuracy_scale=2.0,
                                                config=gwn.QueryConfig(sort_queries=0)).raise_if_error()
    torch.cuda.synchronize()
    assert np.array_equal(_bits(w[~over]), _bits(ww.cpu().numpy()[~over]))

    d, _, flags = _run(gwn, point, dev, 2, q, cap=1, bvh=bvh, a=float("inf"))
    over = flags.astype(bool)
    assert over.any() and np.array_equal(np.isnan(d), over)

    from tests import mesh_e2e

    rays = [torch.from_numpy(a).to(dev) for a in mesh_e2e.fixture_rays(mesh)]
    hits, _, flags = _run(gwn, point, dev, 3, rays, cap=1, bvh=bvh, a=0.0, b=float("inf"), rays=True)
    assert np.array_equal(hits["status"] == gwn.RayStatus.OVERFLOW, flags.astype(bool))
    assert flags.any()

    soup = synth.open_soup(32, 16)
    geom, bvh, _, chain = _objects(gwn, soup, chain=True)
    w, _, flags = _run(gwn, point, dev, 4, q, cap=1, geom=geom, bvh=bvh, chain=chain)
    over = flags.astype(bool)
    assert over.any() and np.isnan(w[over]).all()


def test_point_queries_on_a_leaf_root(gwn, orc, point, dev):
    """One triangle: the BVH is a single leaf and the point functions take the leaf-root paths."""
    import torch

    one = synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), [[0, 1, 2]])
    geom, bvh, mom, _ = _objects(gwn, one, order=2)
    q = [torch.tensor(v, device=dev) for v in ([0.2, -1.0, 0.3], [0.3, 0.5, 0.1], [0.5, -2.0, 0.7])]
    w, _, flags = _run(gwn, point, dev, 1, q, order=2, bvh=bvh, mom=mom, a=2.0)
    ww = torch.empty(3, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ww, accuracy_scale=2.0).raise_if_error()
    e, _, _ = _run(gwn, point, dev, 0, q, bvh=bvh)
    d, _, _ = _run(gwn, point, dev, 2, q, bvh=bvh, a=float("inf"))
    torch.cuda.synchronize()
    assert not flags.any()
    assert np.array_equal(_bits(w), _bits(ww.cpu().numpy()))
    assert np.abs(e - orc.winding_exact_mesh_f64(one, *[t.cpu().numpy() for t in q])).max() <= 2e-6
    ref = orc.unsigned_distance_f32(orc.build_bvh4(one), *[t.cpu().numpy() for t in q])
    assert np.array_equal(_bits(d), _bits(ref))
