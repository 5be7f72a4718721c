"""GPU parity of the Taylor gradient batch (gwn_compute_winding_gradient_taylor_batch).

The device traversal takes the same accept/descend decisions as the oracle's replay (work
counters equal query by query) and sums the same terms in the same order.  Values differ only by
FP32 rounding: the device contracts products into FMAs and uses the hardware reciprocal square
root, the replay does neither, so the gradient is compared with a relative tolerance.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES, build_gpu

pytestmark = pytest.mark.gpu

OPEN = synth.open_soup(64, 32, seed=11)  # open mesh: the winding number varies off the surface


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _moment(gwn, bvh, order):
    mom = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    return mom


def _run(gwn, dev, bvh, mom, qx, qy, qz, beta, precision, sort, with_out=True):
    import torch

    n = qx.shape[0]
    d = [_t(a, dev) for a in (qx, qy, qz)]
    g = [torch.full((n,), float("nan"), dtype=torch.float32, device=dev) for _ in range(3)]
    out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev) if with_out else None
    cnt = torch.zeros(3 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_gradient_taylor_batch(
        bvh, mom, *d, *g, accuracy_scale=beta, precision=precision, out=out, counters=cnt,
        config=gwn.QueryConfig(sort_queries=sort)).raise_if_error()
    torch.cuda.synchronize()
    return (out.cpu().numpy() if with_out else None), [v.cpu().numpy() for v in g], \
        cnt.cpu().numpy().reshape(-1, 3).astype(np.uint64)


@pytest.mark.parametrize("sort", [0, 1], ids=["per_lane", "wave"])
@pytest.mark.parametrize("precision", ["plain", "split"])
@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("beta", [0.0, 2.0, 4.0])
def test_gradient_matches_oracle_replay(gwn, orc, dev, order, beta, precision, sort):
    mesh = OPEN
    _, bvh = build_gpu(gwn, mesh)
    mom = _moment(gwn, bvh, order)
    qx, qy, qz = synth.mixed_queries(mesh, 8192, seed=31, band=0.01)
    split = precision == "split"
    w, g, cnt = _run(gwn, dev, bvh, mom, qx, qy, qz, beta,
                     gwn.Precision.SPLIT if split else gwn.Precision.PLAIN, sort)
    rb = orc.build_bvh4(mesh)
    rw, *rg, rc = orc.winding_gradient_taylor_f32(rb, orc.refit_moment(rb, order), order, qx, qy, qz,
                                                  accuracy_scale=beta, counters=True, near_fp64=split)
    assert np.array_equal(cnt, rc), "identical accept/descend decisions"
    # Same terms, same order: the values differ only by FP32 rounding inside the terms (device FMA
    # contraction and hardware rsqrt).  Near an edge the Biot-Savart terms of adjacent triangles
    # cancel, which amplifies that rounding; the bound therefore adds a multiple of the FP32 error
    # the replay itself shows against its FP64-near-field form on the same traversal.
    if split:
        env_w, env_g = 0.0, [0.0, 0.0, 0.0]
    else:
        sw, *sg = orc.winding_gradient_taylor_f32(rb, orc.refit_moment(rb, order), order, qx, qy, qz,
                                                  accuracy_scale=beta, near_fp64=True)
        env_w = 4 * np.abs(rw.astype(np.float64) - sw)
        env_g = [4 * np.abs(rg[k].astype(np.float64) - sg[k]) for k in range(3)]
    dw = np.abs(w.astype(np.float64) - rw)
    assert (dw <= 1e-5 * (1 + np.abs(rw)) + env_w).all(), f"winding: max rel {(dw / (1 + np.abs(rw))).max():.3e}"
    for k in range(3):
        dg = np.abs(g[k].astype(np.float64) - rg[k])
        # Far-field terms are FP32 in both precision modes, so both share the far-field bound.
        ok = dg <= 1e-4 * (1 + np.abs(rg[k])) + env_g[k]
        assert ok.all(), f"axis {k}: {int((~ok).sum())} queries, max rel {(dg / (1 + np.abs(rg[k]))).max():.3e}"


def test_gradient_only_output_and_winding_agree(gwn, dev):
    """The gradient-only entry point gives the same gradient as the combined form (same kernel),
    and the combined form's winding (the reference's expression on the traversal records) agrees
    with the winding batch's (the same expansion folded into the evaluation records) to rounding:
    identical decisions, terms differing by a few ulps."""
    import torch

    _, bvh = build_gpu(gwn, OPEN)
    mom = _moment(gwn, bvh, 2)
    qx, qy, qz = synth.mixed_queries(OPEN, 4096, seed=5, band=0.01)
    _, g1, _ = _run(gwn, dev, bvh, mom, qx, qy, qz, 2.0, gwn.Precision.PLAIN, 0, with_out=False)
    w2, g2, _ = _run(gwn, dev, bvh, mom, qx, qy, qz, 2.0, gwn.Precision.PLAIN, 0)
    for k in range(3):
        assert np.array_equal(g1[k], g2[k])
    d = [_t(a, dev) for a in (qx, qy, qz)]
    w = torch.empty(4096, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, w, accuracy_scale=2.0,
                                                config=gwn.QueryConfig(sort_queries=0)).raise_if_error()
    w = w.cpu().numpy()
    rel = (np.abs(w - w2) / (1 + np.abs(w2))).max()
    print(f"[gradient] combined vs value-only winding: max rel {rel:.3e}", flush=True)
    assert rel <= 2e-6


def test_descended_leaf_gradient_matches_exact(gwn, orc, dev):
    """tests/unit_gwn_query_gradient.cu:211-240 on an open mesh: beta = 1e6 descends every
    leaf; the float gradient is within 3e-3 of the FP64 exact gradient, the split form within 1e-5."""
    _, bvh = build_gpu(gwn, OPEN)
    mom = _moment(gwn, bvh, 2)
    qx, qy, qz = synth.near_surface_log_queries(OPEN, 512, seed=2, lo=1e-4, hi=1e-1)
    ref = orc.winding_gradient_exact_mesh_f64(OPEN, qx, qy, qz)
    _, g, _ = _run(gwn, dev, bvh, mom, qx, qy, qz, 1e6, gwn.Precision.PLAIN, 0)
    _, gs, _ = _run(gwn, dev, bvh, mom, qx, qy, qz, 1e6, gwn.Precision.SPLIT, 1)
    for k in range(3):
        scale = 1 + np.abs(ref[k])
        assert (np.abs(g[k] - ref[k]) / scale).max() <= 3e-3
        assert (np.abs(gs[k] - ref[k]) / scale).max() <= 1e-5


def test_far_field_gradient_order_progression(gwn, orc, dev):
    """tests/unit_gwn_query_gradient.cu:153-199 (beta = 0, the reference's octahedron queries)."""
    mesh = MESHES["asym_octahedron"]
    _, bvh = build_gpu(gwn, mesh)
    qx = np.float32([2.9, -2.7, 0.4, -0.6, 3.2, -2.8, 1.8, -2.1])
    qy = np.float32([0.4, -0.8, 3.1, -2.9, 2.2, 1.7, -2.6, 2.4])
    qz = np.float32([1.1, -0.5, 0.8, -1.2, -2.3, 2.5, 2.8, -2.5])
    ref = np.stack(orc.winding_gradient_exact_mesh_f64(mesh, qx, qy, qz), 1)
    errs = []
    for order in (0, 1, 2):
        _, g, _ = _run(gwn, dev, bvh, _moment(gwn, bvh, order), qx, qy, qz, 0.0, gwn.Precision.PLAIN, 0)
        errs.append(np.linalg.norm(np.stack(g, 1) - ref, axis=1).mean())
    assert errs[1] < errs[0] and errs[2] < errs[1] and errs[2] < 3e-3


def test_gradient_nan_query_and_leaf_root(gwn, orc, dev):
    """tests/unit_gwn_query_gradient.cu:269-288; a one-triangle (leaf-root) BVH."""
    _, bvh = build_gpu(gwn, MESHES["asym_octahedron"])
    mom = _moment(gwn, bvh, 2)
    _, g, _ = _run(gwn, dev, bvh, mom, np.float32([np.nan]), np.float32([0.25]), np.float32([3.0]), 0.0,
                   gwn.Precision.PLAIN, 0)
    assert all(np.isnan(v[0]) for v in g)
    single = MESHES["single"]
    _, bvh1 = build_gpu(gwn, single)
    mom1 = _moment(gwn, bvh1, 1)
    qx, qy, qz = np.float32([0.2, -1.0]), np.float32([0.3, 0.5]), np.float32([0.5, -2.0])
    w, g, _ = _run(gwn, dev, bvh1, mom1, qx, qy, qz, 2.0, gwn.Precision.PLAIN, 0)
    ref = orc.winding_gradient_exact_mesh_f64(single, qx, qy, qz)
    for k in range(3):
        assert np.abs(g[k] - ref[k]).max() <= 1e-5


def test_gradient_validation_contract(gwn, dev):
    """gwn_query.cuh:711-737: invalid BVH, stale moment, negative scale, span mismatches."""
    import torch

    _, bvh = build_gpu(gwn, MESHES["torus"])
    mom = _moment(gwn, bvh, 1)
    q = [torch.zeros(4, device=dev) for _ in range(3)]
    g = [torch.zeros(4, device=dev) for _ in range(3)]
    st = gwn.gwn_compute_winding_gradient_taylor_batch(gwn.BvhObject(), mom, *q, *g)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "no queryable data" in st.message
    st = gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *q, *g, accuracy_scale=-1.0)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "nonnegative" in st.message
    st = gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *q, g[0], g[1], torch.zeros(3, device=dev))
    assert st.error == gwn.Error.INVALID_ARGUMENT
    st = gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *q, *g,
                                                       config=gwn.QueryConfig(stack_capacity=1))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Traversal stack capacity" in st.message
    other = gwn.BvhObject()
    gwn.gwn_build_bvh(build_gpu(gwn, MESHES["torus"])[0], other,
                      gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    st = gwn.gwn_compute_winding_gradient_taylor_batch(other, mom, *q, *g)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
