"""GPU parity at the five benchmarked workloads (small, large, precision split, batched,
differentiation), through the C ABI against the oracle.

Each test builds its configuration at full size with the seeded generators and checks either all of
it (when the oracle finishes in seconds) or a strided subsample, plus size-independent properties
of the full run:

- small (icosphere 4x, 65,536 queries and rays, LBVH, Taylor order 1, scale 2): everything;
- large (1,048,576-triangle displaced sphere, H-PLOC radius 8, 16,777,216 mixed queries): the
  headline Taylor order-2 split batch at accuracy scale 6.25 in Morton order, and orders 1 and 2
  plain at scale 2, on 16,384 strided queries: the per-query work counters equal the oracle's
  replay, values agree to its tolerance, and max |dW| <= 1e-4 against FP64 (the GPU's FP64 exact
  form on all 16,384, itself within 1e-10 of the CPU FP64 oracle on 1,024 of them);
- precision split (262K-triangle open soup, 4,194,304 log-offset queries): Taylor order 2 split on
  all queries, exact and antipodal (plain and split) on a 262,144-query subsample, against the
  GPU's FP64 exact form, and the oracle's replays on 4,096 queries;
- batched (triangle counts 200 to 200,000, query counts 64 to 1,048,576): one batch bit-identical
  to sequential single-mesh calls;
- differentiation (131,072-triangle sphere, 1,048,576 labelled queries, 262,144 targets): dL/dV of
  the full step, its linearity over query halves, the winding adjoint on 4,096 queries against the
  FP64 oracle adjoint and against FP64 central differences (h = 1e-4) on 64 random vertices, and
  the distance adjoint on 2,048 targets against the FP64 oracle.

Tolerances follow the reference's tests: FP32 GPU vs CPU winding 2e-5 + 2e-5 |w|
(tests/integration_gwn_builder.cu:234-239); Taylor vs the replay of the same traversal 1e-5
relative (2e-6 with the FP64 near field); FP64 forms 1e-10 (tests/integration_gwn_builder.cu:
432-435 uses 1e-12 on meshes a thousand times smaller).
"""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _objects(gwn, mesh, method, radius=8, order=None):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method, hploc_search_radius=radius)).raise_if_error()
    mom = None
    if order is not None:
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    return geom, bvh, mom


def _assert_tree(orc, mesh, bvh, method, radius=8):
    ref = orc.build_bvh4(mesh, method=method, radius=radius)
    nodes, prim, tris = bvh.download()
    assert nodes.tobytes() == ref.nodes.tobytes(), "BVH nodes differ from the oracle"
    assert np.array_equal(prim, ref.primitive_indices)
    assert np.array_equal(tris.view(np.uint32), ref.triangles.view(np.uint32))
    return ref


def _assert_moments(orc, ref, bvh, mom, order):
    want = orc.refit_moment(ref, order)
    got = mom.download(bvh.info()["node_count"])
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "moments differ from the oracle"
    return want


def _taylor(gwn, dev, bvh, mom, d, beta, precision, sort=1, stack=128):
    import torch

    n = d[0].numel()
    out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    cnt = torch.zeros(3 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(
        bvh, mom, *d, out, accuracy_scale=beta, precision=precision, counters=cnt,
        config=gwn.QueryConfig(sort_queries=sort, stack_capacity=stack)).raise_if_error()
    torch.cuda.synchronize()
    return out.cpu().numpy(), cnt.cpu().numpy().reshape(-1, 3).astype(np.uint64)


def _fp64_exact(gwn, dev, bvh, d):
    import torch

    out = torch.empty(d[0].numel(), dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *d, out, gwn.Precision.FP64).raise_if_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


# ---------------------------------------------------------------------------------------------
# Small: icosphere subdivided 4x, 65,536 queries and rays, LBVH, BVH4, Taylor order 1, scale 2.
# ---------------------------------------------------------------------------------------------

def test_small_config_all_outputs(gwn, orc, dev):
    import torch

    print("[config] small", flush=True)
    mesh, q, rays = synth.config_small()
    assert (mesh.vertex_count, mesh.triangle_count, q[0].shape[0], rays[0].shape[0]) == (2562, 5120, 65536, 65536)
    _, bvh, mom = _objects(gwn, mesh, gwn.BuildMethod.LBVH, order=1)
    ref = _assert_tree(orc, mesh, bvh, "lbvh")
    rmom = _assert_moments(orc, ref, bvh, mom, 1)
    d = [_t(a, dev) for a in q]
    n = q[0].shape[0]
    exact = torch.empty(n, dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_exact_batch(bvh, *d, exact).raise_if_error()
    ref64 = orc.winding_exact_mesh_f64(mesh, *q)
    assert np.all(np.abs(exact.cpu().numpy() - ref64) <= 2e-5 + 2e-5 * np.abs(ref64))
    assert np.abs(_fp64_exact(gwn, dev, bvh, d) - ref64).max() <= 1e-12
    for sort in (0, 1):
        got, cnt = _taylor(gwn, dev, bvh, mom, d, 2.0, gwn.Precision.PLAIN, sort=sort, stack=64)
        want, wcnt = orc.winding_taylor_f32(ref, rmom, 1, *q, accuracy_scale=2.0, counters=True)
        assert np.array_equal(cnt, wcnt)
        assert (np.abs(got.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 1e-5
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    gwn.gwn_compute_unsigned_distance_batch(bvh, *d, dist).raise_if_error()
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), orc.unsigned_distance_f32(ref, *q).view(np.uint32))
    hits = torch.empty(8 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_ray_first_hit_batch(bvh, *[_t(a, dev) for a in rays], hits).raise_if_error()
    torch.cuda.synchronize()
    got = hits.cpu().numpy().view(gwn.RAY_HIT_DTYPE)
    assert got.tobytes() == orc.ray_first_hit_f32(ref, *rays).tobytes()
    assert (got["status"] == 1).mean() > 0.5
    # Work counters of the same traversals: an unbounded distance search reaches at least one
    # triangle and tests at least the root's children; a hit ray tested at least one triangle.
    work = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_unsigned_distance_work(bvh, *d, work).raise_if_error()
    torch.cuda.synchronize()
    w = work.cpu().numpy().reshape(-1, 2)
    root_children = int(sum(1 for c in ref.nodes[0]["children"] if int(c["reference"]) >> 47 & 1))
    assert (w[:, 0] >= root_children).all() and (w[:, 1] >= 1).all()
    gwn.gwn_compute_ray_first_hit_work(bvh, *[_t(a, dev) for a in rays], work).raise_if_error()
    torch.cuda.synchronize()
    w = work.cpu().numpy().reshape(-1, 2)
    assert (w[got["status"] == 1, 1] >= 1).all() and (w[:, 0] >= 1).all()


# ---------------------------------------------------------------------------------------------
# Large: 1,048,576-triangle displaced sphere, H-PLOC radius 8, 16,777,216 mixed queries.
# ---------------------------------------------------------------------------------------------

LARGE_QUERIES = 16_777_216
LARGE_SAMPLE = 16_384
HEADLINE_SCALE = 6.25


@pytest.fixture(scope="module")
def large(gwn, orc, dev):
    print("[config] large: mesh, queries, H-PLOC r8 build", flush=True)
    mesh = synth.config_large_mesh()
    q = synth.mixed_queries(mesh, LARGE_QUERIES, near_fraction=0.5, band=0.01)
    idx = np.linspace(0, LARGE_QUERIES - 1, LARGE_SAMPLE).astype(np.int64)
    qs = tuple(np.ascontiguousarray(a[idx]) for a in q)
    del q
    geom, bvh, _ = _objects(gwn, mesh, gwn.BuildMethod.HPLOC, 8)
    ref = _assert_tree(orc, mesh, bvh, "hploc", 8)
    d = [_t(a, dev) for a in qs]
    ref64 = _fp64_exact(gwn, dev, bvh, d)
    cpu64 = orc.winding_exact_mesh_f64(mesh, *(a[CPU_SUB] for a in qs))
    return {"mesh": mesh, "geom": geom, "bvh": bvh, "ref": ref, "q": qs, "d": d, "fp64": ref64, "cpu64": cpu64}


CPU_SUB = slice(None, None, LARGE_SAMPLE // 1024)  # 1,024 queries for the CPU FP64 oracle


def test_large_fp64_exact_matches_cpu_oracle(large):
    """The GPU FP64 exact form the accuracy checks use, against the CPU FP64 oracle (1,024 queries)."""
    assert np.abs(large["fp64"][CPU_SUB] - large["cpu64"]).max() <= 1e-10


def test_large_headline_taylor_order2_split(gwn, orc, dev, large):
    """The benchmarked batch: order 2, FP64 near field, accuracy scale 6.25, Morton order."""
    print("[config] large headline", flush=True)
    bvh, ref, q = large["bvh"], large["ref"], large["q"]
    mom = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    rmom = _assert_moments(orc, ref, bvh, mom, 2)
    stack = max(64, bvh.info()["internal_stack_bound"])
    got, cnt = _taylor(gwn, dev, bvh, mom, large["d"], HEADLINE_SCALE, gwn.Precision.SPLIT, sort=1, stack=stack)
    want, wcnt = orc.winding_taylor_f32(ref, rmom, 2, *q, accuracy_scale=HEADLINE_SCALE, counters=True,
                                        near_fp64=True, stack_capacity=stack)
    assert np.array_equal(cnt, wcnt), "accept/descend decisions differ from the oracle replay"
    assert (np.abs(got.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 2e-6
    err = np.abs(got.astype(np.float64) - large["fp64"])
    print(f"  max |dW| {err.max():.3e}, p99 {np.quantile(err, 0.99):.3e}", flush=True)
    assert err.max() <= 1e-4
    assert np.abs(got[CPU_SUB].astype(np.float64) - large["cpu64"]).max() <= 1e-4


@pytest.mark.parametrize("order", [1, 2])
def test_large_taylor_plain_scale2(gwn, orc, dev, large, order):
    """Orders 1 and 2 at the reference's default accuracy scale, plain FP32, per-lane and wave."""
    print(f"[config] large plain order={order}", flush=True)
    bvh, ref, q = large["bvh"], large["ref"], large["q"]
    mom = gwn.MomentObject(order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    rmom = _assert_moments(orc, ref, bvh, mom, order)
    stack = max(64, bvh.info()["internal_stack_bound"])
    want, wcnt = orc.winding_taylor_f32(ref, rmom, order, *q, accuracy_scale=2.0, counters=True,
                                        stack_capacity=stack)
    for sort in (0, 1):
        got, cnt = _taylor(gwn, dev, bvh, mom, large["d"], 2.0, gwn.Precision.PLAIN, sort=sort, stack=stack)
        assert np.array_equal(cnt, wcnt)
        assert (np.abs(got.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 1e-5
    # The approximation error itself, against FP64 (orders improve monotonically).
    err = np.abs(got.astype(np.float64) - large["fp64"]).max()
    assert err <= {1: 0.1, 2: 0.02}[order]


# ---------------------------------------------------------------------------------------------
# Precision split: 262K-triangle open soup, 4,194,304 queries at log-uniform normal offsets.
# ---------------------------------------------------------------------------------------------

def test_precision_split_config(gwn, orc, dev):
    import torch

    print("[config] precision split: soup and 4M queries", flush=True)
    soup = synth.open_soup(612, 306, seed=synth.DEFAULT_SEED)
    assert 250_000 < soup.triangle_count < 270_000
    n = 4_194_304
    q = synth.near_surface_log_queries(soup, n, seed=synth.DEFAULT_SEED + 5, lo=1e-6, hi=1e-1)
    geom, bvh, mom = _objects(gwn, soup, gwn.BuildMethod.HPLOC, 8, order=2)
    ref = _assert_tree(orc, soup, bvh, "hploc", 8)
    rmom = _assert_moments(orc, ref, bvh, mom, 2)
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain).raise_if_error()
    d = [_t(a, dev) for a in q]
    stack = max(64, bvh.info()["internal_stack_bound"])
    # FP64 reference on a strided subsample; the GPU FP64 form is pinned to the CPU oracle on 2,048.
    idx = np.linspace(0, n - 1, 262_144).astype(np.int64)
    didx = _t(idx, dev)
    sub = [a[didx].contiguous() for a in d]
    qsub = tuple(np.ascontiguousarray(a[idx]) for a in q)
    ref64 = _fp64_exact(gwn, dev, bvh, sub)
    pin = slice(None, None, 128)
    assert np.abs(ref64[pin] - orc.winding_exact_mesh_f64(soup, *(a[pin] for a in qsub))).max() <= 1e-10
    # Taylor order 2 split on all 4M queries.
    full = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, full, accuracy_scale=HEADLINE_SCALE,
                                                precision=gwn.Precision.SPLIT,
                                                config=gwn.QueryConfig(stack_capacity=stack)).raise_if_error()
    torch.cuda.synchronize()
    taylor = full.cpu().numpy()
    assert np.isfinite(taylor).all()
    err = np.abs(taylor[idx].astype(np.float64) - ref64)
    print(f"  taylor o2 split: max |dW| {err.max():.3e} p99 {np.quantile(err, 0.99):.3e}", flush=True)
    assert err.max() <= 1e-4
    rs = slice(None, None, 64)  # 4,096 queries replayed by the oracle
    got, cnt = _taylor(gwn, dev, bvh, mom, [a[rs].contiguous() for a in sub], HEADLINE_SCALE, gwn.Precision.SPLIT,
                       sort=1, stack=stack)
    want, wcnt = orc.winding_taylor_f32(ref, rmom, 2, *(a[rs] for a in qsub), accuracy_scale=HEADLINE_SCALE,
                                        counters=True, near_fp64=True, stack_capacity=stack)
    assert np.array_equal(cnt, wcnt)
    assert (np.abs(got.astype(np.float64) - want) / (1 + np.abs(want))).max() <= 2e-6
    # Exact and antipodal, plain and split, on the subsample.
    m = idx.size
    o32 = torch.empty(m, dtype=torch.float32, device=dev)
    o64 = torch.empty(m, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *sub, o64, gwn.Precision.SPLIT).raise_if_error()
    torch.cuda.synchronize()
    e = np.abs(o64.cpu().numpy() - ref64)
    print(f"  exact split: max |dW| {e.max():.3e}", flush=True)
    assert e.max() <= 1e-6
    gwn.gwn_compute_winding_number_exact_batch(bvh, *sub, o32).raise_if_error()
    torch.cuda.synchronize()
    plain = o32.cpu().numpy()
    e_plain = np.abs(plain - ref64)
    print(f"  exact plain: max |dW| {e_plain.max():.3e}", flush=True)
    assert e_plain.max() <= 0.05 and np.quantile(e_plain, 0.99) <= 1e-3
    # The plain batch is the reference's float arithmetic: equal to the CPU FP32 replay's order of
    # operations up to the device's atan2/sqrt rounding.  Close to the surface the FP32 sums cancel
    # and each side's rounding is amplified on its own, so compare the queries at least 1e-3 from
    # the soup (unsigned distance batch) with the reference's GPU-vs-CPU tolerance.
    r32 = orc.winding_exact_records_f32(ref.triangles, *(a[rs] for a in qsub))
    dist = torch.empty(rs_n := qsub[0][rs].shape[0], dtype=torch.float32, device=dev)
    gwn.gwn_compute_unsigned_distance_batch(bvh, *[a[rs].contiguous() for a in sub], dist,
                                            config=gwn.QueryConfig(stack_capacity=stack)).raise_if_error()
    torch.cuda.synchronize()
    well = dist.cpu().numpy() >= 1e-3
    assert rs_n == 4096 and well.mean() > 0.3
    assert np.all(np.abs(plain[rs][well] - r32[well]) <= 2e-5 + 2e-5 * np.abs(ref64[rs][well]))
    boundary = orc.boundary_chain(soup)
    for precision, bound in ((gwn.Precision.SPLIT, 1e-6), (gwn.Precision.PLAIN, None)):
        gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *sub, o32, precision=precision,
                                                       config=gwn.QueryConfig(stack_capacity=stack)).raise_if_error()
        torch.cuda.synchronize()
        anti = o32.cpu().numpy()
        ok = np.isfinite(anti)
        assert ok.mean() > 0.999
        ea = np.abs(anti[ok] - ref64[ok])
        print(f"  antipodal {precision.name}: max |dW| {ea.max():.3e}, NaN {int((~ok).sum())}", flush=True)
        if bound is not None:
            assert ea.max() <= bound
        want = orc.winding_antipodal_f32(ref, soup, boundary, *(a[rs] for a in qsub),
                                         near_fp64=precision == gwn.Precision.SPLIT, stack_capacity=stack)
        both = np.isfinite(want) & np.isfinite(anti[rs])
        assert np.array_equal(np.isfinite(want), np.isfinite(anti[rs]))
        assert np.abs(anti[rs][both].astype(np.float64) - want[both]).max() <= 1e-5


# ---------------------------------------------------------------------------------------------
# Batched: meshes of 200 to 200,000 triangles, 64 to 1,048,576 queries each.
# ---------------------------------------------------------------------------------------------

def test_batched_config_size_range_bit_identical(gwn, dev):
    """The extremes of the batched distribution (200 and 200,000 triangles; 64 and 1,048,576
    queries) and log-uniform draws between them, tori and displaced spheres under random affine
    maps: batch query and a graph replay with refits equal sequential calls bit for bit."""
    import torch

    print("[config] batched size range", flush=True)
    rng = np.random.default_rng(2024)
    tri_counts = [200, 200_000] + [int(np.exp(rng.uniform(np.log(200), np.log(200_000)))) for _ in range(22)]
    query_counts = [1 << 20, 64] + [int(np.exp(rng.uniform(np.log(64), np.log(1 << 20)))) for _ in range(22)]
    items, refs, keep = [], [], []
    for k, (target, n) in enumerate(zip(tri_counts, query_counts)):
        if k % 2 == 0:
            nu = max(3, int(np.sqrt(target / 2)))
            base = synth.torus(nu, max(3, int(target / (2 * nu))))
        else:
            nlon = max(4, int(np.sqrt(target / 2)))
            base = synth.displaced_sphere(nlon, max(3, int(target / (2 * nlon))), seed=k)
        rot, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        v = (np.stack([base.vx, base.vy, base.vz], 1).astype(np.float64) @ rot.T * rng.uniform(0.5, 2.0)
             + rng.uniform(-5, 5, size=3)).astype(np.float32)
        mesh = synth._mesh_from_arrays(v, np.stack([base.i0, base.i1, base.i2], 1))
        geom, bvh, mom = _objects(gwn, mesh, gwn.BuildMethod.HPLOC, 8, order=1)
        lo, hi = v.min(0), v.max(0)
        c, e = (lo + hi) / 2, (hi - lo) / 2 * 1.5
        q = [_t(a, dev) for a in synth.uniform_box_points(1000 + k, n, c - e, c + e)]
        out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
        ref = torch.empty_like(out)
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *q, ref, accuracy_scale=2.0).raise_if_error()
        items.append((bvh, mom, q[0], q[1], q[2], out))
        refs.append(ref)
        keep.append(geom)
    assert min(b.info()["triangle_count"] for b, *_ in items) <= 220
    assert max(b.info()["triangle_count"] for b, *_ in items) >= 190_000
    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    batch.tune()
    batch.query().raise_if_error()
    torch.cuda.synchronize()

    def check():
        for k, ((*_, out), ref) in enumerate(zip(items, refs)):
            got, want = out.cpu().numpy(), ref.cpu().numpy()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"mesh {k} differs"

    check()
    batch.capture(include_refit=True).raise_if_error()
    for it in items:
        it[5].fill_(float("nan"))
    batch.replay().raise_if_error()
    torch.cuda.synchronize()
    check()
    assert batch.status().is_ok()


# ---------------------------------------------------------------------------------------------
# Differentiation: 131,072-triangle sphere, 1,048,576 labelled queries, 262,144 targets.
# ---------------------------------------------------------------------------------------------

EPS32 = float(np.finfo(np.float32).eps)


def test_differentiation_config_gradients(gwn, orc, dev):
    import torch

    from smallgwn_amd import autograd as ag

    print("[config] differentiation: full step", flush=True)
    mesh = synth.displaced_sphere(256, 256)
    assert mesh.triangle_count == 131_072
    faces = np.stack([mesh.i0, mesh.i1, mesh.i2], 1)
    verts = torch.tensor(np.stack([mesh.vx, mesh.vy, mesh.vz], 1), device=dev, requires_grad=True)
    scene = ag.MeshScene(faces, verts, order=1, method=gwn.BuildMethod.HPLOC, accuracy_scale=2.0)
    qx, qy, qz = synth.uniform_box_points(21, 1 << 20, (-1.5,) * 3, (1.5,) * 3)
    queries = torch.tensor(np.stack([qx, qy, qz], 1), device=dev)
    labels = (queries.norm(dim=1) < 0.9).to(torch.float32)
    tx, ty, tz = synth.uniform_box_points(22, 1 << 18, (-1.0,) * 3, (1.0,) * 3)
    targets = torch.tensor(np.stack([tx, ty, tz], 1), device=dev)
    targets = 0.9 * targets / targets.norm(dim=1, keepdim=True)

    w = ag.taylor_winding(verts, queries, scene)
    w_leaf = w.detach().requires_grad_(True)
    lw = ((torch.sigmoid(20.0 * (w_leaf - 0.5)) - labels) ** 2).mean()
    (u_full,) = torch.autograd.grad(lw, w_leaf)
    dist = ag.unsigned_distance(verts, targets, scene)
    loss = ((torch.sigmoid(20.0 * (w - 0.5)) - labels) ** 2).mean() + dist.mean()
    loss.backward()
    g_full = verts.grad.detach().double().cpu().numpy()
    assert np.isfinite(g_full).all() and np.abs(g_full).max() > 0

    geom, bvh, mom = scene.geometry, scene.bvh, scene.moment
    V = mesh.vertex_count

    def winding_adjoint(idx):
        dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
        qd = [queries[idx, k].contiguous() for k in range(3)]
        gwn.gwn_compute_winding_taylor_adjoint_batch(geom, bvh, mom, *qd, u_full[idx].contiguous(), *dv,
                                                     accuracy_scale=2.0).raise_if_error()
        torch.cuda.synchronize()
        return np.stack([a.double().cpu().numpy() for a in dv])

    # Linearity over the two query halves (each above the wave-cooperative threshold).
    n = queries.shape[0]
    all_idx = torch.arange(n, device=dev)
    g_all = winding_adjoint(all_idx)
    g_halves = winding_adjoint(all_idx[: n // 2]) + winding_adjoint(all_idx[n // 2:])
    # The two sides re-associate FP32 atomic sums of per-query terms that largely cancel on the
    # closed sphere, so the bound is relative to the largest vertex total elementwise and tighter
    # in the L2 norm.
    scale = np.abs(g_all).max()
    assert np.abs(g_all - g_halves).max() <= 2e-3 * scale
    assert np.linalg.norm(g_all - g_halves) <= 2e-4 * np.linalg.norm(g_all)

    # Winding adjoint on 4,096 strided queries against the FP64 oracle (magnitude envelope) ...
    print("[config] differentiation: oracle adjoint and central differences", flush=True)
    sidx = np.linspace(0, n - 1, 4096).astype(np.int64)
    g_sub = winding_adjoint(torch.from_numpy(sidx).to(dev))
    qs = [np.ascontiguousarray(a[sidx]) for a in (qx, qy, qz)]
    us = u_full[torch.from_numpy(sidx).to(dev)].double().cpu().numpy()
    rb = orc.build_bvh4(mesh, method="hploc", radius=8)
    rm = orc.refit_moment(rb, 1)
    rdv, _ = orc.winding_taylor_adjoint_f64(rb, rm, 1, 2.0, mesh, *qs, us)
    mv, _ = orc.adjoint_magnitude_f64(mesh, *qs, us)
    bound = 1e-4 * np.abs(rdv).max() + 128.0 * EPS32 * mv
    assert (np.abs(g_sub - rdv) <= bound).all()
    rel = np.linalg.norm(g_sub - rdv) / np.linalg.norm(rdv)
    print(f"  winding dL/dV vs FP64 oracle: relative error {rel:.2e}", flush=True)
    assert rel <= 1e-3
    # ... which central differences of the FP64 frozen sum pin at h = 1e-4 on 64 random vertices.
    rng = np.random.default_rng(64)
    picks = rng.choice(V, 64, replace=False)
    base = np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)
    fd = np.zeros((3, 64))
    h = 1e-4
    for j, vtx in enumerate(picks):
        for k in range(3):
            vp, vm = base.copy(), base.copy()
            vp[k, vtx] += h
            vm[k, vtx] -= h
            fp = orc.winding_taylor_frozen_f64(rb, rm, 1, 2.0, mesh, *qs, us, vertices=vp)
            fm = orc.winding_taylor_frozen_f64(rb, rm, 1, 2.0, mesh, *qs, us, vertices=vm)
            fd[k, j] = (fp - fm) / (2 * h)
    ana = rdv[:, picks]
    fd_rel = np.linalg.norm(ana - fd) / np.linalg.norm(fd)
    gpu_rel = np.linalg.norm(g_sub[:, picks] - fd) / np.linalg.norm(fd)
    print(f"  central differences: oracle rel {fd_rel:.2e}, GPU rel {gpu_rel:.2e}", flush=True)
    assert fd_rel <= 1e-4 and gpu_rel <= 1e-3

    # Distance adjoint on 2,048 strided targets against the FP64 brute-force oracle.
    tidx = np.linspace(0, targets.shape[0] - 1, 2048).astype(np.int64)
    tsel = targets[torch.from_numpy(tidx).to(dev)]
    ut = np.full(2048, 1.0 / targets.shape[0], np.float32)
    dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
    gwn.gwn_compute_unsigned_distance_adjoint_batch(geom, bvh, *[tsel[:, k].contiguous() for k in range(3)],
                                                    _t(ut, dev), *dv).raise_if_error()
    torch.cuda.synchronize()
    gd = np.stack([a.double().cpu().numpy() for a in dv])
    tp = [np.ascontiguousarray(tsel[:, k].cpu().numpy()) for k in range(3)]
    _, rdd, _ = orc.distance_adjoint_f64(mesh, *tp, ut)
    # Each target moves the three vertices of its nearest triangle only, so dL/dV is sparse and a
    # vertex total is one or two targets' terms: where two triangles are nearest within FP32
    # rounding, the FP32 search and the FP64 brute force can credit different vertices.  Bound
    # the whole field in the L2 norm and each entry by a percent of the largest.
    rel = np.linalg.norm(gd - rdd) / np.linalg.norm(rdd)
    print(f"  distance dL/dV vs FP64 oracle: relative error {rel:.2e}", flush=True)
    assert rel <= 1e-3
    assert np.abs(gd - rdd).max() <= 1e-2 * np.abs(rdd).max()
