"""Pins the CPU oracle before it is trusted as the checker of the HIP path.

Every check here comes from outside the oracle's own formulas: the reference's
known-answer tests (file:line cited), closed forms (triangle second moments,
octant solid angle, divergence theorem), brute-force definitions (bit-loop Morton
interleave, top-down radix tree, exhaustive union of leaf boxes) or an independent
library (numpy's stable argsort, Gauss-Legendre quadrature).
"""

import math

import numpy as np
import pytest

from smallgwn_amd import synth
from tests import helpers


# ---------------------------------------------------------------------------------------------
# Morton codes
# ---------------------------------------------------------------------------------------------

# Known answers of the reference's quantize-and-interleave (include/gwn/detail/
# gwn_bvh_build_common.cuh:76-88), evaluated from the definition.
MORTON_KATS = [
    ((0.0, 0.0, 0.0), 0x0, 0x0),
    ((1.0, 1.0, 1.0), 0x7FFFFFFFFFFFFFFF, 0x3FFFFFFF),
    ((1.0, 0.0, 0.0), 0x4924924924924924, 0x24924924),
    ((0.0, 1.0, 0.0), 0x2492492492492492, 0x12492492),
    ((0.0, 0.0, 1.0), 0x1249249249249249, 0x09249249),
    ((0.5, 0.5, 0.5), 0x0FFFFFFFFFFFFFFF, 0x07FFFFFF),
    ((0.25, 0.75, 0.125), 0x21BFFFFFFFFFFFFF, 0x10DFFFFF),
]


@pytest.mark.parametrize("point,e64,e32", MORTON_KATS)
def test_morton_known_answers(orc, point, e64, e32):
    assert orc.morton_u64(*point) == e64
    assert orc.morton_u32(*point) == e32


def test_morton_matches_bit_loop_and_clamps(orc):
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.25, 1.25, size=(500, 3)).astype(np.float32)
    for x, y, z in pts:
        q = [int(np.float32(min(max(v, np.float32(0)), np.float32(1))) * np.float32(2097151.0)) for v in (x, y, z)]
        assert orc.morton_u64(float(x), float(y), float(z)) == helpers.morton_bitloop_u64(*q)
        q10 = [int(np.float32(min(max(v, np.float32(0)), np.float32(1))) * np.float32(1023.0)) for v in (x, y, z)]
        assert orc.morton_u32(float(x), float(y), float(z)) == helpers.morton_bitloop_u32(*q10)


# ---------------------------------------------------------------------------------------------
# Stable sort and LBVH
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 2, 257, 5000])
def test_stable_order_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 40, size=n, dtype=np.uint64) << np.uint64(37)  # heavy duplicates
    keys |= rng.integers(0, 3, size=n, dtype=np.uint64)
    assert np.array_equal(orc.stable_order_u64(keys), np.argsort(keys, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lbvh_matches_topdown_radix_tree(orc, seed):
    """Karras' parallel search equals the sequential top-down definition, duplicates included."""
    rng = np.random.default_rng(seed)
    T = 300
    v = rng.uniform(-1, 1, size=(3 * T, 3)).astype(np.float32)
    if seed == 2:
        v[: 3 * T // 2] = v[0]  # many coincident triangles -> equal Morton codes
    f = np.arange(3 * T).reshape(T, 3)
    mesh = synth._mesh_from_arrays(v, f)
    bvh = orc.build_bvh4(mesh)
    ref = helpers.lbvh_topdown(bvh.sorted_morton)
    assert len(ref) == T - 1
    for nid, (l, ll, r, rl) in ref.items():
        got = bvh.binary_nodes[nid]
        assert (int(got[0]), int(got[1])) == (l, r)
        assert (int(got[2]) == 2) == ll and (int(got[3]) == 2) == rl


def _scene_morton_numpy(mesh):
    """Scene inverse and per-triangle codes restated in numpy float32 (build_common.cuh:170-213, 243-274)."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.int64)
    smin, smax = v.min(0), v.max(0)
    ext = np.where(smax == smin, np.float32(1), smax - smin).astype(np.float32)
    inv = np.where(smax == smin, np.float32(1), np.float32(1) / ext).astype(np.float32)
    tri = v[f]
    lo, hi = tri.min(1), tri.max(1)
    c = (lo + hi) * np.float32(0.5)
    nrm = np.clip((c - smin) * inv, np.float32(0), np.float32(1)) * np.float32(2097151.0)
    q = nrm.astype(np.uint64)
    codes = np.array([helpers.morton_bitloop_u64(*map(int, r)) for r in q], dtype=np.uint64)
    return codes


@pytest.mark.parametrize("name", ["icosphere2", "asym_octahedron", "soup", "coincident33", "torus"])
def test_build_bvh4_invariants(orc, name):
    if name == "icosphere2":
        mesh = synth.icosphere(2)
    elif name == "asym_octahedron":
        mesh = synth.asymmetric_octahedron()
    elif name == "soup":
        rng = np.random.default_rng(3)
        v = rng.normal(size=(600, 3)).astype(np.float32)
        mesh = synth._mesh_from_arrays(v, rng.integers(0, 600, size=(400, 3)))
    elif name == "coincident33":
        # tests/unit_gwn_bvh.cu:376-384: 33 copies of one triangle
        v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
        mesh = synth._mesh_from_arrays(v, np.tile([[0, 1, 2]], (33, 1)))
    else:
        mesh = synth.torus(24, 12)
    bvh = orc.build_bvh4(mesh)
    codes = _scene_morton_numpy(mesh)
    assert np.array_equal(bvh.sorted_morton, np.sort(codes, kind="stable"))
    assert np.array_equal(bvh.primitive_indices, np.argsort(codes, kind="stable").astype(np.uint32))
    helpers.check_bvh4(bvh.nodes, bvh.root_reference, bvh.primitive_indices, bvh.triangles, mesh,
                       bvh.internal_stack_bound, bvh.packed_stack_bound)
    helpers.check_collapse_rule(bvh.nodes, bvh.binary_nodes, bvh.binary_bounds)


def test_build_leaf_root_kat(orc):
    """tests/unit_gwn_bvh.cu:93-139: one triangle -> leaf root, bounds (0,0,0)-(1,1,0)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    bvh = orc.build_bvh4(synth._mesh_from_arrays(v, [[0, 1, 2]]))
    assert bvh.node_count == 0
    assert bvh.root_reference == helpers.VALID_BIT | (1 << 48)
    assert bvh.root_bounds.tolist() == [0, 0, 0, 1, 1, 0]
    assert bvh.internal_stack_bound == 0 and bvh.packed_stack_bound == 0
    rec = bvh.triangles[0]
    assert rec.tolist() == [0, 0, 0, 1, 0, 0, 0, 1, 0]
    assert (rec[3:6] - rec[0:3]).tolist() == [1, 0, 0] and (rec[6:9] - rec[0:3]).tolist() == [0, 1, 0]


def test_build_two_triangle_kat(orc):
    """tests/unit_gwn_bvh.cu:141-218: two triangles -> one node with two single-primitive leaves."""
    x = [0, 1, 0, 4, 6, 4]
    y = [0, 0, 2, 0, 0, 3]
    z = [1, 1, 1, -2, -2, -2]
    mesh = synth._mesh_from_arrays(np.stack([x, y, z], 1), [[0, 1, 2], [3, 4, 5]])
    bvh = orc.build_bvh4(mesh)
    assert bvh.node_count == 1
    kids = bvh.nodes[0]["children"]
    leaves = [int(k["reference"]) for k in kids if int(k["reference"]) & helpers.VALID_BIT]
    assert len(leaves) == 2 and all((r >> 48) == 1 for r in leaves)
    assert sorted(bvh.primitive_indices[[r & helpers.OFFSET_MASK for r in leaves]].tolist()) == [0, 1]
    helpers.check_bvh4(bvh.nodes, bvh.root_reference, bvh.primitive_indices, bvh.triangles, mesh, 0, 1)


def _deform(mesh, seed: int):
    """Smooth deformation of a mesh's vertices, same indexing."""
    rng = np.random.default_rng(seed)
    a = rng.normal(size=3).astype(np.float32) * np.float32(0.3)
    vx = (mesh.vx * np.float32(1.2) + a[0] * np.sin(np.float32(2.0) * mesh.vy)).astype(np.float32)
    vy = (mesh.vy * np.float32(0.8) + a[1] * np.cos(np.float32(3.0) * mesh.vz)).astype(np.float32)
    vz = (mesh.vz + a[2] * mesh.vx * mesh.vy).astype(np.float32)
    return synth.Mesh(vx, vy, vz, mesh.i0, mesh.i1, mesh.i2)


def test_refit_bvh_kat(orc):
    """tests/unit_gwn_bvh.cu:220-287: moved two-triangle mesh -> root (-7,-4,-6)-(11,9,3),
    primitive order preserved, records taken from the new positions."""
    x = [0, 1, 0, 4, 6, 4]
    y = [0, 0, 2, 0, 0, 3]
    z = [1, 1, 1, -2, -2, -2]
    faces = [[0, 1, 2], [3, 4, 5]]
    bvh = orc.build_bvh4(synth._mesh_from_arrays(np.stack([x, y, z], 1), faces))
    moved = synth._mesh_from_arrays(
        np.stack([[-7, -2, -7, 8, 11, 8], [5, 5, 9, -4, -4, 2], [3, 3, 3, -6, -6, -6]], 1), faces)
    out = orc.refit_bvh(bvh, moved)
    assert out.root_bounds.tolist() == [-7, -4, -6, 11, 9, 3]
    assert np.array_equal(out.primitive_indices, bvh.primitive_indices)
    assert out.internal_stack_bound == bvh.internal_stack_bound
    helpers.check_bvh4(out.nodes, out.root_reference, out.primitive_indices, out.triangles, moved, 0, 1)


def test_refit_bvh_leaf_root(orc):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    bvh = orc.build_bvh4(synth._mesh_from_arrays(v, [[0, 1, 2]]))
    out = orc.refit_bvh(bvh, synth._mesh_from_arrays(v * 3 - 1, [[0, 1, 2]]))
    assert out.root_bounds.tolist() == [-1, -1, -1, 2, 2, -1]
    assert out.triangles[0].tolist() == (v * 3 - 1).reshape(-1).tolist()


@pytest.mark.parametrize("name", ["torus", "icosphere3", "soup"])
def test_refit_bvh_unchanged_is_identity_and_deformed_is_exact(orc, name):
    if name == "torus":
        mesh = synth.torus(24, 12)
    elif name == "icosphere3":
        mesh = synth.icosphere(3)
    else:
        rng = np.random.default_rng(5)
        v = rng.normal(size=(500, 3)).astype(np.float32)
        mesh = synth._mesh_from_arrays(v, rng.integers(0, 500, size=(700, 3)))
    bvh = orc.build_bvh4(mesh)
    same = orc.refit_bvh(bvh, mesh)
    assert same.nodes.tobytes() == bvh.nodes.tobytes()
    assert np.array_equal(same.root_bounds.view(np.uint32), bvh.root_bounds.view(np.uint32))
    moved = _deform(mesh, 7)
    out = orc.refit_bvh(bvh, moved)
    # Child boxes are exact subtree min/max of the new triangles; topology and stack bounds keep.
    helpers.check_bvh4(out.nodes, out.root_reference, out.primitive_indices, out.triangles, moved,
                       bvh.internal_stack_bound, bvh.packed_stack_bound)
    refs = out.nodes["children"]["reference"]
    assert np.array_equal(refs, bvh.nodes["children"]["reference"])
    v = np.stack([moved.vx, moved.vy, moved.vz], 1)
    assert out.root_bounds[:3].tolist() == v.min(0).tolist() or name == "soup"
    assert out.root_bounds[3:].tolist() == v.max(0).tolist() or name == "soup"


# ---------------------------------------------------------------------------------------------
# Moments
# ---------------------------------------------------------------------------------------------

def test_triangle_payload_matches_closed_forms(orc):
    """Order 0/1 are exact for a flat triangle; the HDK split integrals equal A/12 sum d d^T."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        tri = rng.normal(size=(3, 3)).astype(np.float32)
        rec = tri.reshape(9)
        p = orc.triangle_payload(rec, 2)
        a, b, c = tri.astype(np.float64)
        n = 0.5 * np.cross(b - a, c - a)
        area = np.linalg.norm(n)
        P = (a + b + c) / 3.0
        assert p[0] == pytest.approx(area, rel=1e-12)
        assert np.allclose(p[1:4], area * P, rtol=1e-12, atol=1e-14)
        assert np.allclose(p[4:7], n, rtol=1e-12, atol=1e-14)
        assert np.allclose(p[7:16].reshape(3, 3), np.outer(n, P), rtol=1e-12, atol=1e-14)
        M = helpers.triangle_second_moment(a, b, c)
        nhat = n / area
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
        raw = p[16:34].reshape(3, 6)
        for i in range(3):
            for k, (j, l) in enumerate(pairs):
                expect = nhat[i] * M[j, l] + n[i] * P[j] * P[l]
                assert raw[i, k] == pytest.approx(expect, rel=1e-9, abs=1e-12)


def test_moment_divergence_theorem_and_precisions(orc):
    """Closed surface: sum of area-weighted normals is zero; FP32 and FP64 refits agree."""
    mesh = synth.icosphere(3)
    bvh = orc.build_bvh4(mesh)
    for order in (0, 1, 2):
        m32 = orc.refit_moment(bvh, order)
        m64 = orc.refit_moment(bvh, order, dtype=np.float64)
        assert m32.shape == (bvh.node_count, orc.field_count(order), 4)
        root_N = m64[0, 4:7, :].sum(axis=1)
        assert np.abs(root_N).max() < 1e-12
        scale = np.abs(m64).max()
        assert np.abs(m32.astype(np.float64) - m64).max() < 2e-6 * scale


def _single_cluster_error(orc, order, dist):
    mesh = synth.asymmetric_octahedron()
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, order)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    q = (d * dist).astype(np.float32)
    exact = orc.winding_exact_mesh_f64(mesh, q[:, 0], q[:, 1], q[:, 2])
    approx = orc.winding_taylor_f32(bvh, mom, order, q[:, 0], q[:, 1], q[:, 2], accuracy_scale=0.0)
    return float(np.abs(approx.astype(np.float64) - exact).mean())


def test_taylor_order_convergence_rates(orc):
    """Far-field error of order k decays like d^-(k+3): pins every term of each expansion
    (tests/unit_gwn_query_winding.cu:146-189 only checks the order progression)."""
    for order, rate in ((0, 3), (1, 4), (2, 5)):
        e1 = _single_cluster_error(orc, order, 40.0)
        e2 = _single_cluster_error(orc, order, 80.0)
        slope = math.log2(e1 / e2)
        assert abs(slope - rate) < 0.5, (order, slope)


def test_taylor_order_progression_reference_case(orc):
    """tests/unit_gwn_query_winding.cu:146-189: beta = 0, errors shrink with order, order 2 < 2e-3."""
    mesh = synth.asymmetric_octahedron()
    bvh = orc.build_bvh4(mesh)
    qx = np.array([2.9, -2.7, 0.4, -0.6, 3.2, -2.8, 1.8, -2.1], np.float32)
    qy = np.array([0.4, -0.8, 3.1, -2.9, 2.2, 1.7, -2.6, 2.4], np.float32)
    qz = np.array([1.1, -0.5, 0.8, -1.2, -2.3, 2.5, 2.8, -2.5], np.float32)
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    err = []
    for order in (0, 1, 2):
        w = orc.winding_taylor_f32(bvh, orc.refit_moment(bvh, order), order, qx, qy, qz, accuracy_scale=0.0)
        err.append(float(np.abs(w - ref).mean()))
    assert err[1] < err[0] and err[2] < err[1] and err[2] < 2e-3


def test_taylor_large_beta_equals_exact(orc):
    """tests/unit_gwn_bvh_moment.cu:42-98: beta = 1e6 descends every cluster."""
    mesh = synth.octahedron()
    bvh = orc.build_bvh4(mesh)
    qx = np.array([0.13, 3.25], np.float32)
    qy = np.array([-0.21, -2.5], np.float32)
    qz = np.array([0.17, 4.75], np.float32)
    exact = orc.winding_exact_records_f32(bvh.triangles, qx, qy, qz)
    for order in (0, 1, 2):
        w = orc.winding_taylor_f32(bvh, orc.refit_moment(bvh, order), order, qx, qy, qz, accuracy_scale=1e6)
        assert np.allclose(w, exact, atol=1e-6)


# ---------------------------------------------------------------------------------------------
# Solid angle and exact winding
# ---------------------------------------------------------------------------------------------

def test_solid_angle_octant_and_quadrature(orc):
    a, b, c = [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]
    o = [0.0, 0.0, 0.0]
    # The octant triangle subtends 1/8 of the sphere; its normal points away from the origin.
    assert orc.solid_angle_f64(a, b, c, o) == pytest.approx(4 * math.pi / 8, rel=1e-14)
    assert orc.solid_angle_f64(a, c, b, o) == pytest.approx(-4 * math.pi / 8, rel=1e-14)
    assert orc.solid_angle_f32(a, b, c, o) == pytest.approx(math.pi / 2, rel=1e-6)
    rng = np.random.default_rng(2)
    for _ in range(20):
        tri = rng.normal(size=(3, 3))
        q = rng.normal(size=3) * 6.0 + 8.0
        assert orc.solid_angle_f64(tri[0], tri[1], tri[2], q) == pytest.approx(
            helpers.solid_angle_quad(tri[0], tri[1], tri[2], q), rel=1e-9, abs=1e-12)
    # Degenerate cases defined by the reference (gwn_query_geometry_impl.cuh:25-39).
    assert orc.solid_angle_f64(a, b, c, a) == 0.0
    flat = ([0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    assert orc.solid_angle_f64(*flat, [0.25, 0.25, 0.0]) == 0.0  # exactly coplanar: numerator 0
    # Just above the face on its normal side: the solid angle tends to -2*pi (w -> -1/2).
    assert orc.solid_angle_f64(*flat, [0.25, 0.25, 1e-9]) == pytest.approx(-2 * math.pi, rel=1e-6)
    assert orc.solid_angle_f64(*flat, [0.25, 0.25, -1e-9]) == pytest.approx(2 * math.pi, rel=1e-6)


@pytest.mark.parametrize("meshname", ["cube", "octahedron", "icosphere3"])
def test_exact_winding_closed_meshes(orc, meshname):
    mesh = {"cube": synth.closed_cube, "octahedron": synth.octahedron, "icosphere3": lambda: synth.icosphere(3)}[meshname]()
    inside = np.array([[0.0, 0.0, 0.0], [0.1, -0.2, 0.15], [-0.3, 0.2, 0.1]], np.float32)
    outside = np.array([[3.0, 0.0, 0.0], [0.0, -2.5, 1.0], [5.0, 5.0, 5.0]], np.float32)
    q = np.concatenate([inside, outside])
    w = orc.winding_exact_mesh_f64(mesh, q[:, 0], q[:, 1], q[:, 2])
    assert np.allclose(w, [1, 1, 1, 0, 0, 0], atol=1e-12)
    bvh = orc.build_bvh4(mesh)
    w32 = orc.winding_exact_records_f32(bvh.triangles, q[:, 0], q[:, 1], q[:, 2])
    assert np.allclose(w32, [1, 1, 1, 0, 0, 0], atol=2e-6)


def test_exact_fp32_matches_fp64_reference_case(orc):
    """tests/unit_gwn_query_winding.cu:193-215: FP32 exact within 2e-6 of the CPU reference."""
    mesh = synth.asymmetric_octahedron()
    bvh = orc.build_bvh4(mesh)
    qx = np.array([0, 4, -2, 0.2, 1.5], np.float32)
    qy = np.array([0, 3, 1, -1.4, 0.2], np.float32)
    qz = np.array([0, 2, -3, 2.2, -1.8], np.float32)
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    w = orc.winding_exact_records_f32(bvh.triangles, qx, qy, qz)
    assert np.abs(w - ref).max() <= 2e-6


def test_nan_query_propagates(orc):
    """tests/unit_gwn_query_winding.cu:222-243."""
    mesh = synth.asymmetric_octahedron()
    bvh = orc.build_bvh4(mesh)
    qx, qy, qz = np.array([np.nan], np.float32), np.array([0.25], np.float32), np.array([3.0], np.float32)
    assert np.isnan(orc.winding_exact_records_f32(bvh.triangles, qx, qy, qz)[0])
    assert np.isnan(orc.winding_taylor_f32(bvh, orc.refit_moment(bvh, 2), 2, qx, qy, qz, accuracy_scale=0.0)[0])


def test_split_replay_descending_everything_equals_fp64_exact(orc):
    """With beta = 1e6 every leaf is evaluated: the split replay (FP64 near triangles, FP32 far
    triangles, FP64 accumulation) must equal the all-FP64 exact sum to FP32-term accuracy, on
    queries from 1e-6 to 1e-1 off the surface."""
    mesh = synth.displaced_sphere(64, 32, seed=5)
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, 2)
    qx, qy, qz = synth.near_surface_log_queries(mesh, 512, seed=9, lo=1e-6, hi=1e-1)
    split = orc.winding_taylor_f32(bvh, mom, 2, qx, qy, qz, accuracy_scale=1e6, near_fp64=True)
    plain = orc.winding_taylor_f32(bvh, mom, 2, qx, qy, qz, accuracy_scale=1e6)
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    assert np.abs(split - ref).max() <= 2e-7
    # The plain FP32 form is measurably worse near the surface (the split is not a no-op).
    assert np.abs(plain - ref).max() > 4 * np.abs(split - ref).max()
