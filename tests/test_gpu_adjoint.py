"""GPU adjoint batches against the FP64 adjoint oracle (which central differences pin).

The device re-traverses with the forward's decisions and accumulates FP32 terms with atomics.
Near the surface the terms are large (~1/d^2) and cancel (on a closed mesh the exact vertex
gradient is zero), so the FP32 result differs from the FP64 one by the rounding of its terms, not
of its value.  Each entry is therefore checked against
    1e-4 * max|ref| + ENV * eps32 * (sum of |term| over all triangles and queries),
the magnitude sum taken in FP64 by the oracle (brute force, every term exact).
"""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _objects(gwn, mesh, order):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    mom = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    return geom, bvh, mom


EPS32 = float(np.finfo(np.float32).eps)
ENV = 128.0  # ulps of the magnitude sum


def _check_winding_adjoint(orc, mesh, q, u, dv, dq, rdv, rdq, calls=1):
    mv, mq = orc.adjoint_magnitude_f64(mesh, *q, u)
    scale = np.abs(rdv).max()
    ev, eq = np.abs(dv - rdv), np.abs(dq - rdq)
    bound_v = 1e-4 * scale + ENV * EPS32 * calls * mv
    bound_q = 1e-4 * (1 + np.abs(rdq)) + ENV * EPS32 * calls * mq
    rv = (ev / (EPS32 * calls * mv))[mv > 0].max(initial=0.0)
    rq = (eq / (EPS32 * calls * mq))[mq > 0].max(initial=0.0)
    print(f"  dV: max err {ev.max():.3e} of {scale:.3e}, max err/(eps mag) {rv:.2f}; "
          f"dq: max err/(eps mag) {rq:.2f}", flush=True)
    assert (ev <= bound_v).all(), f"dV: {int((ev > bound_v).sum())} entries beyond the bound"
    assert (eq <= bound_q).all(), f"dq: {int((eq > bound_q).sum())} entries beyond the bound"


def _winding_adjoint(gwn, dev, geom, bvh, mom, q, u, beta):
    import torch

    V = geom.vertex_count()
    dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
    dq = [torch.zeros(q[0].shape[0], dtype=torch.float32, device=dev) for _ in range(3)]
    gwn.gwn_compute_winding_taylor_adjoint_batch(geom, bvh, mom, *[_t(a, dev) for a in q], _t(u, dev), *dv, *dq,
                                                 accuracy_scale=beta).raise_if_error()
    torch.cuda.synchronize()
    return np.stack([a.cpu().numpy() for a in dv]), np.stack([a.cpu().numpy() for a in dq], 1)


@pytest.mark.parametrize("beta", [2.0, 1e6])
@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("name", ["open_soup", "displaced"])
def test_winding_adjoint_matches_oracle(gwn, orc, dev, name, order, beta):
    mesh = synth.open_soup(24, 12, seed=3) if name == "open_soup" else synth.displaced_sphere(48, 24)
    print(f"[adjoint] winding {name} order={order} beta={beta}", flush=True)
    geom, bvh, mom = _objects(gwn, mesh, order)
    q = synth.mixed_queries(mesh, 4096, seed=9, band=0.02)
    u = np.random.default_rng(order).normal(size=4096).astype(np.float32)
    dv, dq = _winding_adjoint(gwn, dev, geom, bvh, mom, q, u, beta)
    rb = orc.build_bvh4(mesh)
    rdv, rdq = orc.winding_taylor_adjoint_f64(rb, orc.refit_moment(rb, order), order, beta, mesh, *q, u)
    _check_winding_adjoint(orc, mesh, q, u, dv, dq, rdv, rdq)


@pytest.mark.parametrize("name,order", [("displaced", 1), ("open_soup", 0)])
def test_winding_adjoint_large_batch_and_accumulation(gwn, orc, dev, name, order):
    """Above 65,536 queries the batch visits them in Morton order with the wave-cooperative kernel
    (slot and triangle adjoints summed over the wave before one atomic); two calls add their
    gradients."""
    import torch

    mesh = synth.displaced_sphere(64, 32) if name == "displaced" else synth.open_soup(32, 16, seed=7)
    print(f"[adjoint] large batch {name} order={order}", flush=True)
    geom, bvh, mom = _objects(gwn, mesh, order)
    q = synth.mixed_queries(mesh, 70000, seed=4, band=0.02)
    u = np.random.default_rng(5).normal(size=70000).astype(np.float32)
    V = geom.vertex_count()
    dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
    dq = [torch.zeros(70000, dtype=torch.float32, device=dev) for _ in range(3)]
    args = [_t(a, dev) for a in q] + [_t(u, dev)]
    for _ in range(2):
        gwn.gwn_compute_winding_taylor_adjoint_batch(geom, bvh, mom, *args, *dv, *dq).raise_if_error()
    torch.cuda.synchronize()
    got = np.stack([a.cpu().numpy() for a in dv])
    rb = orc.build_bvh4(mesh)
    rdv, rdq = orc.winding_taylor_adjoint_f64(rb, orc.refit_moment(rb, order), order, 2.0, mesh, *q, u)
    # dq is written (not added) by each call; dV adds both calls' terms.
    _check_winding_adjoint(orc, mesh, q, u, got, np.stack([a.cpu().numpy() for a in dq], 1), 2 * rdv, rdq, calls=2)


def test_distance_adjoint_matches_oracle(gwn, orc, dev):
    import torch

    mesh = synth.open_soup(24, 12, seed=3)
    geom, bvh, _ = _objects(gwn, mesh, 0)
    p = synth.uniform_box_points(11, 4096, (-1.4,) * 3, (1.4,) * 3)
    u = np.random.default_rng(3).normal(size=4096).astype(np.float32)
    V = geom.vertex_count()
    dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
    dp = [torch.zeros(4096, dtype=torch.float32, device=dev) for _ in range(3)]
    gwn.gwn_compute_unsigned_distance_adjoint_batch(geom, bvh, *[_t(a, dev) for a in p], _t(u, dev), *dv,
                                                    *dp).raise_if_error()
    torch.cuda.synchronize()
    _, rdv, rdp = orc.distance_adjoint_f64(mesh, *p, u)
    got = np.stack([a.cpu().numpy() for a in dv])
    assert np.abs(got - rdv).max() <= 1e-4 * np.abs(rdv).max()
    assert np.abs(np.stack([a.cpu().numpy() for a in dp], 1) - rdp).max() <= 1e-4 * (1 + np.abs(rdp).max())


def test_adjoint_leaf_root_and_validation(gwn, orc, dev):
    import torch

    one = synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), [[0, 1, 2]])
    geom, bvh, mom = _objects(gwn, one, 1)
    q = (np.float32([0.2, -1.0, 0.3]), np.float32([0.3, 0.5, 0.1]), np.float32([0.5, -2.0, 0.7]))
    u = np.float32([1.0, -0.5, 2.0])
    dv, dq = _winding_adjoint(gwn, dev, geom, bvh, mom, q, u, 2.0)
    rb = orc.build_bvh4(one)
    rdv, rdq = orc.winding_taylor_adjoint_f64(rb, orc.refit_moment(rb, 1), 1, 2.0, one, *q, u.astype(np.float64))
    assert np.abs(dv - rdv).max() <= 1e-4 * np.abs(rdv).max()
    # Order 2 moments are rejected; vertex gradient spans must match the vertex count.
    geom2, bvh2, mom2 = _objects(gwn, one, 2)
    z = [torch.zeros(3, dtype=torch.float32, device=dev) for _ in range(3)]
    st = gwn.gwn_compute_winding_taylor_adjoint_batch(geom2, bvh2, mom2, *[_t(a, dev) for a in q], _t(u, dev), *z)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "orders 0 and 1" in st.message
    st = gwn.gwn_compute_winding_taylor_adjoint_batch(geom, bvh, mom, *[_t(a, dev) for a in q], _t(u, dev),
                                                      z[0], z[1], z[2][:2])
    assert st.error == gwn.Error.INVALID_ARGUMENT


def test_distance_adjoint_from_forward_records(gwn, orc, dev):
    """The forward's nearest records (gwn_compute_unsigned_distance_batch_ex_f32) name a triangle at
    exactly the returned distance, and the adjoint from them (no traversal) equals the re-traversal
    adjoint: dL/dp bit for bit, dL/dV up to atomic summation order.  Morton-ordered (70,000 points)
    and input-ordered batches; a culling band makes some points record nothing."""
    import torch

    mesh = synth.open_soup(24, 12, seed=3)
    geom, bvh, _ = _objects(gwn, mesh, 0)
    _, _, tris = bvh.download()
    tris = np.asarray(tris, np.float32).reshape(-1, 9)
    V = geom.vertex_count()
    for n, band in ((4096, float("inf")), (70000, 0.3)):
        p = synth.uniform_box_points(13, n, (-1.4,) * 3, (1.4,) * 3)
        u = np.random.default_rng(n).normal(size=n).astype(np.float32)
        d = [_t(a, dev) for a in p]
        dist = torch.empty(n, dtype=torch.float32, device=dev)
        plain = torch.empty_like(dist)
        near = torch.empty(n, dtype=torch.int32, device=dev)
        gwn.gwn_compute_unsigned_distance_batch(bvh, *d, dist, culling_band=band, nearest=near).raise_if_error()
        gwn.gwn_compute_unsigned_distance_batch(bvh, *d, plain, culling_band=band).raise_if_error()
        outs = []
        for use_records in (False, True):
            dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
            dp = [torch.zeros(n, dtype=torch.float32, device=dev) for _ in range(3)]
            gwn.gwn_compute_unsigned_distance_adjoint_batch(
                geom, bvh, *d, _t(u, dev), *dv, *dp, culling_band=band,
                nearest=near if use_records else None).raise_if_error()
            torch.cuda.synchronize()
            outs.append((np.stack([a.cpu().numpy() for a in dv]), np.stack([a.cpu().numpy() for a in dp])))
        dist, plain, rec = dist.cpu().numpy(), plain.cpu().numpy(), near.cpu().numpy().view(np.uint32)
        assert np.array_equal(dist.view(np.uint32), plain.view(np.uint32))
        valid = rec < tris.shape[0]
        assert valid.any() and (~valid).any() == np.isfinite(band)
        for i in np.flatnonzero(valid)[:256]:
            t = tris[rec[i]]
            a, e1, e2 = t[0:3], t[3:6] - t[0:3], t[6:9] - t[0:3]
            d2 = orc.point_triangle_distance2_f32(np.float32([p[0][i], p[1][i], p[2][i]]), a, a + e1, a + e2)
            assert np.float32(np.sqrt(np.float32(d2))) == dist[i]
        (v0, p0), (v1, p1) = outs
        assert np.array_equal(p0.view(np.uint32), p1.view(np.uint32))
        assert np.abs(v0 - v1).max() <= 1e-5 * max(1.0, np.abs(v0).max())
