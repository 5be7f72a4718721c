"""GPU parity of the winding-number batches against the oracle.

Tolerances: FP32 exact within 2e-6 of the FP64 reference (tests/unit_gwn_query_winding.cu:
213-214); FP64 exact within 1e-12 (tests/integration_gwn_builder.cu:432-435); Taylor within
1e-5 of the oracle's FP32 replay of the same traversal (the device uses the hardware
reciprocal square root where the replay divides by sqrt, and sums in the same order).
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES, build_gpu

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _queries(n, seed=1, box=1.5):
    return synth.uniform_box_points(seed, n, (-box,) * 3, (box,) * 3)


def test_exact_reference_case(gwn, orc, dev):
    """tests/unit_gwn_query_winding.cu:193-215."""
    import torch

    mesh = MESHES["asym_octahedron"]
    _, bvh = build_gpu(gwn, mesh)
    qx = np.array([0, 4, -2, 0.2, 1.5], np.float32)
    qy = np.array([0, 3, 1, -1.4, 0.2], np.float32)
    qz = np.array([0, 2, -3, 2.2, -1.8], np.float32)
    out = torch.empty(5, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(bvh, _t(qx, dev), _t(qy, dev), _t(qz, dev), out).raise_if_error()
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    assert np.abs(out.cpu().numpy() - ref).max() <= 2e-6


@pytest.mark.parametrize("name", ["icosphere4", "soup", "torus"])
def test_exact_precisions_match_oracle(gwn, orc, dev, name):
    import torch

    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    qx, qy, qz = _queries(4096, seed=7)
    d = [_t(a, dev) for a in (qx, qy, qz)]
    plain = torch.empty(4096, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(bvh, *d, plain).raise_if_error()
    f64 = torch.empty(4096, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *d, f64, gwn.Precision.FP64).raise_if_error()
    split = torch.empty(4096, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *d, split, gwn.Precision.SPLIT).raise_if_error()
    ref64 = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    ref32 = orc.winding_exact_records_f32(orc.build_bvh4(mesh).triangles, qx, qy, qz)
    assert np.abs(f64.cpu().numpy() - ref64).max() <= 1e-12
    # Reference tolerance for float GPU vs CPU winding: 2e-5 + 2e-5 * |w|
    # (tests/integration_gwn_builder.cu:234-239).
    tol = 2e-5 + 2e-5 * np.abs(ref64)
    assert np.all(np.abs(plain.cpu().numpy() - ref32) <= tol)
    assert np.all(np.abs(plain.cpu().numpy() - ref64) <= tol)
    # The split keeps only far-field terms in FP32: ten times tighter.
    assert np.all(np.abs(split.cpu().numpy() - ref64) <= 0.1 * tol)


def test_exact_nan_and_empty(gwn, dev):
    import torch

    _, bvh = build_gpu(gwn, MESHES["asym_octahedron"])
    q = [_t(np.array([v], np.float32), dev) for v in (np.nan, 0.25, 3.0)]
    out = torch.empty(1, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(bvh, *q, out).raise_if_error()
    assert np.isnan(out.cpu().numpy()[0])
    e = torch.empty(0, dtype=torch.float32, device=dev)
    assert gwn.gwn_compute_winding_number_exact_batch(bvh, e, e, e, e).is_ok()


def _moment(gwn, bvh, order):
    m = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, m).raise_if_error()
    return m


def _assert_structures_match_oracle(orc, mesh, bvh, mom, order):
    """Before comparing query values, pin the tree and the coefficients they are computed from."""
    ref_bvh = orc.build_bvh4(mesh)
    nodes, prim, _ = bvh.download()
    assert nodes.tobytes() == ref_bvh.nodes.tobytes(), "BVH nodes differ from the oracle"
    assert np.array_equal(prim, ref_bvh.primitive_indices)
    got = mom.download(bvh.info()["node_count"])
    assert np.array_equal(got.view(np.uint32), orc.refit_moment(ref_bvh, order).view(np.uint32)), \
        "moments differ from the oracle"
    return ref_bvh


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("name", ["icosphere4", "displaced_256x128", "soup"])
@pytest.mark.parametrize("beta", [0.0, 1.0, 2.0, 4.0])
@pytest.mark.parametrize("sort", [0, 1], ids=["per_lane", "wave"])
def test_taylor_matches_oracle_replay(gwn, orc, dev, name, order, beta, sort):
    import torch

    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    mom = _moment(gwn, bvh, order)
    if beta == 0.0:
        # beta = 0 accepts every non-singular child, so (as in tests/unit_gwn_query_winding.cu:
        # 146-189) it is exercised with far-field queries; near a cluster centroid its terms
        # grow like |q - P|^-3 and only relative agreement is meaningful.
        d = np.stack(_queries(8192, seed=11), axis=1).astype(np.float64)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r = 3.0 + 3.0 * synth.uniform01(17, 8192)
        qx, qy, qz = (np.ascontiguousarray((d * r[:, None])[:, k], dtype=np.float32) for k in range(3))
    else:
        qx, qy, qz = _queries(8192, seed=11)
    out = torch.full((8192,), float("nan"), dtype=torch.float32, device=dev)  # unwritten -> NaN
    cnt = torch.zeros(3 * 8192, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, _t(qx, dev), _t(qy, dev), _t(qz, dev), out,
                                                accuracy_scale=beta, counters=cnt,
                                                config=gwn.QueryConfig(sort_queries=sort)).raise_if_error()
    ref_bvh = _assert_structures_match_oracle(orc, mesh, bvh, mom, order)
    ref, ref_cnt = orc.winding_taylor_f32(ref_bvh, orc.refit_moment(ref_bvh, order), order, qx, qy, qz,
                                          accuracy_scale=beta, counters=True)
    got = out.cpu().numpy()
    err = np.abs(got.astype(np.float64) - ref)
    rel = err / (1.0 + np.abs(ref))
    assert rel.max() <= 1e-5
    # Identical accept/descend decisions: the work counters match query by query.
    assert np.array_equal(cnt.cpu().numpy().reshape(-1, 3).astype(np.uint64), ref_cnt)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("sort", [0, 1], ids=["per_lane", "wave"])
def test_taylor_split_matches_oracle_replay(gwn, orc, dev, order, sort):
    """FP64 near field: same decisions as the oracle's split replay, FP64-accurate leaf sums."""
    import torch

    mesh = MESHES["displaced_256x128"]
    _, bvh = build_gpu(gwn, mesh)
    mom = _moment(gwn, bvh, order)
    qx, qy, qz = synth.mixed_queries(mesh, 8192, seed=21, band=0.01)
    out = torch.full((8192,), float("nan"), dtype=torch.float32, device=dev)  # unwritten -> NaN
    cnt = torch.zeros(3 * 8192, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, _t(qx, dev), _t(qy, dev), _t(qz, dev), out,
                                                accuracy_scale=4.0, precision=gwn.Precision.SPLIT,
                                                counters=cnt, config=gwn.QueryConfig(sort_queries=sort)).raise_if_error()
    ref_bvh = _assert_structures_match_oracle(orc, mesh, bvh, mom, order)
    ref, ref_cnt = orc.winding_taylor_f32(ref_bvh, orc.refit_moment(ref_bvh, order), order, qx, qy, qz,
                                          accuracy_scale=4.0, counters=True, near_fp64=True)
    rel = np.abs(out.cpu().numpy().astype(np.float64) - ref) / (1.0 + np.abs(ref))
    assert rel.max() <= 2e-6
    assert np.array_equal(cnt.cpu().numpy().reshape(-1, 3).astype(np.uint64), ref_cnt)


def test_query_order_is_the_hilbert_order(gwn, dev):
    """The presort visits queries in the order of their Hilbert keys on the 1024^3 grid over the
    root box, ties in input order (a stable sort): equal to numpy's stable argsort of the keys the
    test-side restatement computes."""
    import torch

    from tests import helpers

    _, bvh = build_gpu(gwn, MESHES["displaced_256x128"])
    qx, qy, qz = _queries(100_003, seed=41, box=1.7)  # some outside the root box: clamped
    d = [_t(a, dev) for a in (qx, qy, qz)]
    order = torch.empty(100_003, dtype=torch.int32, device=dev)
    gwn.gwn_compute_query_order(bvh, *d, order).raise_if_error()
    cells = helpers.query_cells(qx, qy, qz, bvh.info()["root_bounds"])
    want = np.argsort(helpers.hilbert_key10(*cells), kind="stable")
    assert np.array_equal(order.cpu().numpy(), want)


def test_query_order_is_a_permutation_and_reusable(gwn, dev):
    import torch

    _, bvh = build_gpu(gwn, MESHES["displaced_256x128"])
    mom = _moment(gwn, bvh, 1)
    qx, qy, qz = _queries(70_000, seed=23)
    d = [_t(a, dev) for a in (qx, qy, qz)]
    order = torch.empty(70_000, dtype=torch.int32, device=dev)
    gwn.gwn_compute_query_order(bvh, *d, order).raise_if_error()
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(70_000))
    a = torch.empty(70_000, dtype=torch.float32, device=dev)
    b = torch.empty(70_000, dtype=torch.float32, device=dev)
    ca = torch.zeros(3 * 70_000, dtype=torch.int32, device=dev)
    cb = torch.zeros(3 * 70_000, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, a, order=order, counters=ca).raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, b, config=gwn.QueryConfig(sort_queries=0),
                                                counters=cb).raise_if_error()
    # Wave-cooperative (ordered) and per-lane (unordered) traversals take identical decisions and
    # sum the same terms in the same order; only FMA contraction may differ between the kernels.
    assert torch.equal(ca, cb)
    rel = ((a - b).abs() / (1 + b.abs())).max().item()
    assert rel <= 1e-6


def test_taylor_presort_does_not_change_results(gwn, dev):
    import torch

    _, bvh = build_gpu(gwn, MESHES["displaced_256x128"])
    mom = _moment(gwn, bvh, 2)
    qx, qy, qz = _queries(100_000, seed=13)
    d = [_t(a, dev) for a in (qx, qy, qz)]
    a = torch.empty(100_000, dtype=torch.float32, device=dev)
    b = torch.empty(100_000, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, a, config=gwn.QueryConfig(sort_queries=0)).raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, b, config=gwn.QueryConfig(sort_queries=1)).raise_if_error()
    # Results land at the input positions whatever the visiting order (ulp-level kernel differences).
    assert ((a - b).abs() / (1 + b.abs())).max().item() <= 1e-6


def test_taylor_split_precision_against_fp64(gwn, orc, dev):
    """FP64 near field: near-surface queries agree with the all-FP64 oracle far better than plain."""
    import torch

    mesh = MESHES["displaced_256x128"]
    _, bvh = build_gpu(gwn, mesh)
    mom = _moment(gwn, bvh, 2)
    qx, qy, qz = synth.near_surface_log_queries(mesh, 2048, seed=3, lo=1e-6, hi=1e-2)
    d = [_t(a, dev) for a in (qx, qy, qz)]
    split = torch.empty(2048, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, split, accuracy_scale=1e6,
                                                precision=gwn.Precision.SPLIT).raise_if_error()
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    # beta = 1e6 descends everything: the split result is the FP64 exact sum rounded to float.
    assert np.abs(split.cpu().numpy() - ref).max() <= 1e-6


def test_taylor_large_beta_equals_exact_and_leaf_root(gwn, dev):
    """tests/unit_gwn_bvh_moment.cu:42-98 and :158-209."""
    import torch

    for name in ("asym_octahedron", "single"):
        _, bvh = build_gpu(gwn, MESHES[name])
        mom = _moment(gwn, bvh, 2)
        q = [_t(np.array(v, np.float32), dev) for v in ([0.13, 3.25, 2.0], [-0.21, -2.5, 2.0], [0.17, 4.75, 2.0])]
        t = torch.empty(3, dtype=torch.float32, device=dev)
        e = torch.empty(3, dtype=torch.float32, device=dev)
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, t, accuracy_scale=1e6 if name != "single" else 0.0).raise_if_error()
        gwn.gwn_compute_winding_number_exact_batch(bvh, *q, e).raise_if_error()
        assert np.allclose(t.cpu().numpy(), e.cpu().numpy(), atol=1e-6)


def test_taylor_validation_contract(gwn, dev):
    """tests/unit_gwn_query_stack.cu:71-76 and tests/unit_gwn_bvh_moment.cu stale-moment checks."""
    import torch

    geom, bvh = build_gpu(gwn, MESHES["icosphere4"])
    mom = _moment(gwn, bvh, 1)
    q = torch.zeros(4, dtype=torch.float32, device=dev)
    out = torch.empty(4, dtype=torch.float32, device=dev)
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, q, q, q, out, config=gwn.QueryConfig(stack_capacity=1))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Traversal stack capacity" in st.message
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, q, q, q, out, accuracy_scale=-1.0)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "accuracy scale" in st.message
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, q, q, q, out, accuracy_scale=float("nan"))
    assert st.error == gwn.Error.INVALID_ARGUMENT
    # A rebuild gives the BVH a new revision: the old moment is stale.
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.LBVH)).raise_if_error()
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, q, q, q, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "not aligned" in st.message
    st = gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, q, q, q[:3], out)
    assert st.error == gwn.Error.INVALID_ARGUMENT


def test_taylor_nan_query(gwn, dev):
    import torch

    _, bvh = build_gpu(gwn, MESHES["asym_octahedron"])
    mom = _moment(gwn, bvh, 2)
    q = [_t(np.array([v], np.float32), dev) for v in (np.nan, 0.25, 3.0)]
    out = torch.empty(1, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, out, accuracy_scale=0.0).raise_if_error()
    assert np.isnan(out.cpu().numpy()[0])


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4096, 4097, 8192, 8193, 70_000, 262_145])
def test_query_order_permutation_repeated(gwn, dev, n):
    """The query presort is a permutation on every call (each query is visited exactly once)."""
    import torch

    _, bvh = build_gpu(gwn, MESHES["icosphere4"])
    qx, qy, qz = _queries(n, seed=n)
    d = [_t(a, dev) for a in (qx, qy, qz)]
    ref = None
    for _ in range(8):
        order = torch.full((n,), -1, dtype=torch.int32, device=dev)
        gwn.gwn_compute_query_order(bvh, *d, order).raise_if_error()
        o = order.cpu().numpy()
        assert np.array_equal(np.sort(o), np.arange(n))
        if ref is None:
            ref = o
        assert np.array_equal(o, ref), "the stable sort is deterministic"
