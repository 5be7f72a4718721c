"""GPU parity of the antipodal batches (gwn_compute_winding_number_antipodal_batch,
gwn_compute_winding_gradient_antipodal_batch).

The device classifies edges and counts crossings with the oracle's unfused FP32 arithmetic over
the same BVH, so singular axes (NaN results included) and crossing counts agree query by query.
The plain winding number differs from the oracle only through atan2f (the oracle rounds a double
atan2), so it is compared within 2e-5 plus 2e-6 relative; the gradient has no transcendental and
the split forms sum FP64 terms, so those are compared within a few FP32 ulps.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _objects(gwn, mesh):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain).raise_if_error()
    return geom, bvh, chain


def _gpu(gwn, dev, mesh, qx, qy, qz, precision, objs=None):
    import torch

    geom, bvh, chain = objs or _objects(gwn, mesh)
    n = qx.shape[0]
    d = [_t(np.float32(a), dev) for a in (qx, qy, qz)]
    w = torch.full((n,), 7.0, dtype=torch.float32, device=dev)
    g = [torch.full((n,), 7.0, dtype=torch.float32, device=dev) for _ in range(3)]
    gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *d, w, precision=precision).raise_if_error()
    gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *d, *g, precision=precision).raise_if_error()
    torch.cuda.synchronize()
    return w.cpu().numpy(), [v.cpu().numpy() for v in g]


def _oracle(orc, mesh, qx, qy, qz, split):
    return orc.winding_antipodal_f32(orc.build_bvh4(mesh), mesh, orc.boundary_chain(mesh), np.float32(qx),
                                     np.float32(qy), np.float32(qz), gradient=True, near_fp64=split,
                                     stack_capacity=128)


def _queries(mesh, n, seed):
    """Uniform points in the padded box plus points on vertices, edge midpoints and face planes,
    where axes become singular."""
    lo = np.array([mesh.vx.min(), mesh.vy.min(), mesh.vz.min()], np.float64)
    hi = np.array([mesh.vx.max(), mesh.vy.max(), mesh.vz.max()], np.float64)
    pad = 0.25 * (hi - lo).max()
    qx, qy, qz = synth.uniform_box_points(seed, n, lo - pad, hi + pad)
    rng = np.random.default_rng(seed)
    tri = rng.integers(0, mesh.i0.shape[0], size=min(64, mesh.i0.shape[0]))
    a = np.stack([mesh.vx[mesh.i0[tri]], mesh.vy[mesh.i0[tri]], mesh.vz[mesh.i0[tri]]], 1)
    b = np.stack([mesh.vx[mesh.i1[tri]], mesh.vy[mesh.i1[tri]], mesh.vz[mesh.i1[tri]]], 1)
    special = np.concatenate([a, np.float32(0.5) * (a + b), a + np.float32([0.0, 0.0, 0.37])]).astype(np.float32)
    return (np.concatenate([qx, special[:, 0]]).astype(np.float32),
            np.concatenate([qy, special[:, 1]]).astype(np.float32),
            np.concatenate([qz, special[:, 2]]).astype(np.float32))


CASES = {
    "open_cube": synth.open_cube,
    "closed_cube": synth.closed_cube,
    "open_soup": lambda: synth.open_soup(48, 24, seed=5),
    "torus": lambda: MESHES["torus"],
    "coincident33": lambda: MESHES["coincident33"],
    "soup": lambda: MESHES["soup"],
}


@pytest.mark.parametrize("split", [False, True], ids=["plain", "split"])
@pytest.mark.parametrize("name", list(CASES))
def test_antipodal_matches_oracle(gwn, orc, dev, name, split):
    mesh = CASES[name]()
    qx, qy, qz = _queries(mesh, 2048, seed=17)
    prec = gwn.Precision.SPLIT if split else gwn.Precision.PLAIN
    w, g = _gpu(gwn, dev, mesh, qx, qy, qz, prec)
    rw, *rg = _oracle(orc, mesh, qx, qy, qz, split)
    print(f"[antipodal] {name} split={split} n={qx.shape[0]} nan={int(np.isnan(rw).sum())}", flush=True)
    assert np.array_equal(np.isnan(w), np.isnan(rw))
    ok = ~np.isnan(rw)
    # Retried X / Y axes carry the reference's +2 pi branch terms, so values reach the hundreds:
    # the bound is relative (a few FP32 ulps of the sum) plus an absolute floor.
    np.testing.assert_allclose(w[ok], rw[ok], rtol=2.5e-7 if split else 2e-6, atol=1e-6 if split else 2e-5)
    # The integer crossing part agrees exactly: on a closed mesh the whole value is an integer.
    if name in ("closed_cube", "torus"):
        assert np.array_equal(w[ok], rw[ok])
    for gv, rv in zip(g, rg):
        assert np.array_equal(np.isnan(gv), np.isnan(rv))
        okg = ~np.isnan(rv)
        scale = max(1.0, float(np.abs(rv[okg]).max())) if okg.any() else 1.0
        np.testing.assert_allclose(gv[okg], rv[okg], rtol=0, atol=4e-6 * scale)


REF_QUERIES = (np.float32([0.23, 2.2, -0.41, -2.3]), np.float32([0.17, 0.3, 0.26, -0.4]),
               np.float32([0.11, 0.2, 1.7, 0.5]))
COLLINEAR_QUERIES = (np.float32([2, 2, -2, -2]), np.float32([-1, 1, -1, 1]), np.float32([2, 2, 2, 2]))


@pytest.mark.parametrize("case", ["cube", "open_cube", "open_cube_collinear"])
def test_antipodal_matches_exact_winding_reference_cases(gwn, orc, dev, case):
    """tests/unit_gwn_query_antipodal.cu:330-396: the closed and open cube at the reference's four
    queries, and the open cube at queries whose boundary projections are collinear, within the
    reference's 3e-4 of the exact winding number."""
    mesh = synth.closed_cube() if case == "cube" else synth.open_cube()
    q = COLLINEAR_QUERIES if case.endswith("collinear") else REF_QUERIES
    ref = orc.winding_exact_mesh_f64(mesh, *q)
    for prec in (gwn.Precision.PLAIN, gwn.Precision.SPLIT):
        w, _ = _gpu(gwn, dev, mesh, *q, prec)
        assert np.abs(w - ref).max() <= 3e-4


def test_antipodal_matches_exact_winding(gwn, orc, dev):
    """A random open soup at 4096 uniform queries: the split form is within 1e-5 of the exact
    winding number.  The plain form is the reference's FP32 arithmetic; near a boundary edge its
    spherical-area terms lose digits, so it is held to the oracle's FP32 replay error plus 2e-5."""
    mesh = synth.open_soup(48, 24, seed=5)
    qx, qy, qz = synth.uniform_box_points(9, 4096, (-1.4,) * 3, (1.4,) * 3)
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    w, _ = _gpu(gwn, dev, mesh, qx, qy, qz, gwn.Precision.PLAIN)
    ws, _ = _gpu(gwn, dev, mesh, qx, qy, qz, gwn.Precision.SPLIT)
    rw = orc.winding_antipodal_f32(orc.build_bvh4(mesh), mesh, orc.boundary_chain(mesh), qx, qy, qz,
                                   stack_capacity=128)
    assert np.all(np.abs(w - ref) <= np.abs(rw.astype(np.float64) - ref) + 2e-5)
    assert np.abs(ws - ref).max() <= 1e-5


def test_antipodal_gradient_matches_exact_gradient(gwn, orc, dev):
    """tests/unit_gwn_query_antipodal.cu:398-438: the boundary-term gradient equals the gradient
    of the exact winding number away from the surface."""
    mesh = synth.open_cube()
    q = (np.float32([0.21, 1.3, -1.4, 0.35]), np.float32([0.18, 0.25, -0.33, 1.4]),
         np.float32([0.15, 0.2, 0.42, 1.55]))
    _, g = _gpu(gwn, dev, mesh, *q, gwn.Precision.SPLIT)
    ref = orc.winding_gradient_exact_mesh_f64(mesh, *q)
    for gv, rv in zip(g, ref):
        np.testing.assert_allclose(gv, rv, rtol=0, atol=1e-5)


def test_antipodal_singular_everywhere_is_nan(gwn, dev):
    w, g = _gpu(gwn, dev, synth.open_cube(), np.float32([1.0]), np.float32([1.0]), np.float32([1.0]),
                gwn.Precision.PLAIN)
    assert np.isnan(w[0]) and all(np.isnan(v[0]) for v in g)


def test_antipodal_validation(gwn, dev):
    import torch

    mesh = synth.open_cube()
    geom, bvh, chain = _objects(gwn, mesh)
    q = [torch.zeros(4, dtype=torch.float32, device=dev) for _ in range(3)]
    out = torch.zeros(4, dtype=torch.float32, device=dev)
    g = [torch.zeros(4, dtype=torch.float32, device=dev) for _ in range(3)]
    st = gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, gwn.BoundaryChainObject(), *q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "unbuilt" in st.message
    other = synth.closed_cube()
    g2, b2, c2 = _objects(gwn, other)
    st = gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, c2, *q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "different mesh shape" in st.message
    big = synth.open_soup(16, 8, seed=1)
    _, bvh_big, _ = _objects(gwn, big)
    st = gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh_big, chain, *q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "triangle counts differ" in st.message
    st = gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *q, out[:3])
    assert st.error == gwn.Error.INVALID_ARGUMENT and "differs from query count" in st.message
    st = gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *q, g[0], g[1], g[2][:2])
    assert st.error == gwn.Error.INVALID_ARGUMENT and "differ from query count" in st.message
    st = gwn.gwn_compute_winding_gradient_antipodal_batch(geom, c2, *q, *g)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "different mesh shape" in st.message
    empty = [torch.zeros(0, dtype=torch.float32, device=dev) for _ in range(4)]
    gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *empty[:3], empty[3]).raise_if_error()
