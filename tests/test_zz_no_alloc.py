"""The allocation-free Taylor batch against the allocating one, through the C ABI.

gwn_compute_winding_number_taylor_batch_noalloc_f32 keeps the Morton presort's scratch in the
library's static per-GPU array.  Two calls in a row (the second smaller, so it runs over the first
call's leftover scratch) must equal gwn_compute_winding_number_taylor_batch_ex_f32 bit for bit in
both precision modes, and a batch above the static capacity is refused before any launch.
"""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu


def _run(gwn, dev, bvh, mom, q, precision, noalloc):
    import torch

    d = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in q]
    out = torch.full((q[0].shape[0],), float("nan"), dtype=torch.float32, device=dev)
    cfg = gwn.QueryConfig(sort_queries=1, stack_capacity=max(64, bvh.info()["internal_stack_bound"]))
    fn = (gwn.gwn_compute_winding_number_taylor_batch_noalloc if noalloc
          else gwn.gwn_compute_winding_number_taylor_batch)
    fn(bvh, mom, *d, out, accuracy_scale=2.0, config=cfg, precision=precision).raise_if_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_taylor_equals_allocating_batch(gwn, dev):
    print("[noalloc] displaced sphere 131k, order 2, 1M then 70k queries", flush=True)
    mesh = synth.displaced_sphere(256, 256)
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=8)).raise_if_error()
    mom = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    for n, seed in ((gwn.NOALLOC_MAX_QUERIES, 31), (70_001, 32)):
        q = synth.uniform_box_points(seed, n, (-1.5,) * 3, (1.5,) * 3)
        for precision in (gwn.Precision.PLAIN, gwn.Precision.SPLIT):
            want = _run(gwn, dev, bvh, mom, q, precision, noalloc=False)
            got = _run(gwn, dev, bvh, mom, q, precision, noalloc=True)
            assert np.isfinite(got).all()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, precision)


def test_noalloc_taylor_refuses_oversize_batch(gwn, dev):
    import torch

    mesh = synth.icosphere(2)
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    mom = gwn.MomentObject(1)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    n = gwn.NOALLOC_MAX_QUERIES + 1
    d = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
    out = torch.full((n,), 7.0, dtype=torch.float32, device=dev)
    st = gwn.gwn_compute_winding_number_taylor_batch_noalloc(bvh, mom, *d, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "static scratch capacity" in st.message
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
