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


# ---------------------------------------------------------------------------------------------
# Gradients
# ---------------------------------------------------------------------------------------------

def _open_mesh():
    """Open displaced sphere (30 % of faces deleted): its winding number varies smoothly off the
    surface, so gradient errors are visible.  On a closed mesh the winding number is piecewise
    constant and the exact gradient vanishes, which would hide a wrong sign or factor."""
    return synth.open_soup(16, 8, seed=11)


def _shell_queries(n, radius, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    q = d / np.linalg.norm(d, axis=1, keepdims=True) * radius
    q = np.round(q * 256) / 256  # dyadic, so q +- 2^-10 is exact in float32
    return tuple(q[:, k].astype(np.float32) for k in range(3))


def _fd_exact_gradient(orc, mesh, qx, qy, qz, h):
    """Central differences of the FP64 exact winding number."""
    out = []
    for axis in range(3):
        p = [np.array(v, dtype=np.float64) for v in (qx, qy, qz)]
        m = [np.array(v, dtype=np.float64) for v in (qx, qy, qz)]
        p[axis] = p[axis] + h
        m[axis] = m[axis] - h
        wp = orc.winding_exact_mesh_f64(mesh, *[v.astype(np.float32) for v in p])
        wm = orc.winding_exact_mesh_f64(mesh, *[v.astype(np.float32) for v in m])
        out.append((wp - wm) / (2 * h))
    return out


def test_exact_gradient_matches_finite_differences(orc):
    """Biot-Savart triangle gradients summed in FP64 = central differences of the FP64 exact
    winding number (h = 2^-10) on an open mesh, inside and outside."""
    mesh = _open_mesh()
    for radius, seed in ((0.5, 1), (1.7, 2)):
        qx, qy, qz = _shell_queries(48, radius, seed)
        g = orc.winding_gradient_exact_mesh_f64(mesh, qx, qy, qz)
        fd = _fd_exact_gradient(orc, mesh, qx, qy, qz, 2.0 ** -10)
        scale = max(np.abs(f).max() for f in fd)
        assert scale > 1e-2  # the open mesh gives a nonvanishing field
        for k in range(3):
            assert np.abs(g[k] - fd[k]).max() <= 1e-5 * scale


def test_closed_mesh_gradient_vanishes(orc):
    """Off the surface of a closed mesh the winding number is constant: zero exact gradient."""
    mesh = synth.asymmetric_octahedron()
    qx, qy, qz = _shell_queries(16, 2.5, 3)
    g = orc.winding_gradient_exact_mesh_f64(mesh, qx, qy, qz)
    assert max(np.abs(v).max() for v in g) < 1e-9


def test_taylor_gradient_descending_everything_equals_exact(orc):
    """tests/unit_gwn_query_gradient.cu:211-240 on an open mesh: beta = 1e6 descends every leaf;
    the replay's reference-form float gradient equals the FP64 exact gradient to 3e-3, the split
    form to FP32-term accuracy."""
    mesh = _open_mesh()
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, 2)
    qx, qy, qz = synth.near_surface_log_queries(mesh, 64, seed=8, lo=1e-3, hi=1e-1)
    ref = orc.winding_gradient_exact_mesh_f64(mesh, qx, qy, qz)
    _, *g = orc.winding_gradient_taylor_f32(bvh, mom, 2, qx, qy, qz, accuracy_scale=1e6)
    _, *gs = orc.winding_gradient_taylor_f32(bvh, mom, 2, qx, qy, qz, accuracy_scale=1e6, near_fp64=True)
    for k in range(3):
        scale = 1 + np.abs(ref[k])
        assert (np.abs(g[k] - ref[k]) / scale).max() <= 3e-3
        assert (np.abs(gs[k] - ref[k]) / scale).max() <= 1e-5


@pytest.mark.parametrize("meshname", ["asym_octahedron", "open"])
def test_taylor_gradient_order_progression(orc, meshname):
    """tests/unit_gwn_query_gradient.cu:153-199 (the reference's closed octahedron and queries),
    and the same on an open mesh: at beta = 0 the mean error vs the exact gradient falls with
    each order; on the octahedron order 2 is below 3e-3."""
    mesh = synth.asymmetric_octahedron() if meshname == "asym_octahedron" else _open_mesh()
    bvh = orc.build_bvh4(mesh)
    if meshname == "asym_octahedron":
        qx = np.float32([2.9, -2.7, 0.4, -0.6, 3.2, -2.8, 1.8, -2.1])
        qy = np.float32([0.4, -0.8, 3.1, -2.9, 2.2, 1.7, -2.6, 2.4])
        qz = np.float32([1.1, -0.5, 0.8, -1.2, -2.3, 2.5, 2.8, -2.5])
    else:
        qx, qy, qz = _shell_queries(16, 4.0, 5)
    ref = np.stack(orc.winding_gradient_exact_mesh_f64(mesh, qx, qy, qz), 1)
    errs = []
    for order in (0, 1, 2):
        _, *g = orc.winding_gradient_taylor_f32(bvh, orc.refit_moment(bvh, order), order, qx, qy, qz,
                                                accuracy_scale=0.0)
        errs.append(np.linalg.norm(np.stack(g, 1) - ref, axis=1).mean())
    assert errs[1] < errs[0] and errs[2] < errs[1]
    if meshname == "asym_octahedron":
        assert errs[2] < 3e-3


@pytest.mark.parametrize("order", [0, 1, 2])
def test_taylor_gradient_is_the_derivative_of_the_taylor_winding(orc, order):
    """At beta = 0 the replay sums the root's expansions, a smooth function of q: its analytic
    gradient must equal central differences of the replayed winding itself (pins every far-field
    gradient term, including the order-1 and order-2 ones, against the winding terms)."""
    mesh = _open_mesh()
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, order)
    qx, qy, qz = _shell_queries(32, 2.5, 6)
    _, *g = orc.winding_gradient_taylor_f32(bvh, mom, order, qx, qy, qz, accuracy_scale=0.0)
    h = 2.0 ** -7
    for axis in range(3):
        p = [qx.copy(), qy.copy(), qz.copy()]
        m = [qx.copy(), qy.copy(), qz.copy()]
        p[axis] += np.float32(h)
        m[axis] -= np.float32(h)
        wp = orc.winding_taylor_f32(bvh, mom, order, *p, accuracy_scale=0.0).astype(np.float64)
        wm = orc.winding_taylor_f32(bvh, mom, order, *m, accuracy_scale=0.0).astype(np.float64)
        fd = (wp - wm) / (2 * h)
        scale = np.abs(g[axis]).max()
        assert scale > 1e-3
        assert np.abs(g[axis] - fd).max() <= 2e-3 * scale + 2e-6


def test_taylor_gradient_nan_query(orc):
    """tests/unit_gwn_query_gradient.cu:269-288: a NaN coordinate gives an all-NaN gradient."""
    mesh = synth.asymmetric_octahedron()
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, 2)
    _, *g = orc.winding_gradient_taylor_f32(bvh, mom, 2, np.float32([np.nan]), np.float32([0.25]),
                                            np.float32([3.0]), accuracy_scale=0.0)
    assert all(np.isnan(v[0]) for v in g)


# ---------------------------------------------------------------------------------------------
# Unsigned distance
# ---------------------------------------------------------------------------------------------

def _segment_distance2(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0) if np.dot(ab, ab) > 0 else 0.0
    d = p - (a + t * ab)
    return float(np.dot(d, d))


def _triangle_distance2_candidates(p, a, b, c):
    """Independent FP64 restatement: the closest point is the plane projection when it falls
    inside the triangle, else the closest point of one of the three edges."""
    n = np.cross(b - a, c - a)
    best = min(_segment_distance2(p, a, b), _segment_distance2(p, b, c), _segment_distance2(p, c, a))
    nn = np.dot(n, n)
    if nn > 0:
        t = np.dot(p - a, n) / nn
        x = p - t * n
        # barycentric sign test
        s1 = np.dot(np.cross(b - a, x - a), n)
        s2 = np.dot(np.cross(c - b, x - b), n)
        s3 = np.dot(np.cross(a - c, x - c), n)
        if s1 >= 0 and s2 >= 0 and s3 >= 0:
            best = min(best, float(np.dot(p - x, p - x)))
    return best


def test_point_triangle_distance_matches_independent_form(orc):
    rng = np.random.default_rng(12)
    for _ in range(400):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2
        got = orc.point_triangle_distance2_f32(p, a, b, c)
        a32, b32, c32, p32 = (v.astype(np.float32).astype(np.float64) for v in (a, b, c, p))
        ref = _triangle_distance2_candidates(p32, a32, b32, c32)
        assert abs(got - ref) <= 1e-5 * (1 + ref)
    # Regions: vertex, edge and face of the unit right triangle.
    a, b, c = np.float32([0, 0, 0]), np.float32([1, 0, 0]), np.float32([0, 1, 0])
    assert orc.point_triangle_distance2_f32(np.float32([-1, -1, 0]), a, b, c) == 2.0
    assert orc.point_triangle_distance2_f32(np.float32([0.5, -2, 0]), a, b, c) == 4.0
    assert orc.point_triangle_distance2_f32(np.float32([0.25, 0.25, 3]), a, b, c) == 9.0


def test_distance_layered_kats(orc):
    """tests/unit_gwn_query_distance_ray.cu:222-270, 350-381 and 555-599."""
    one = orc.build_bvh4(synth.layered_triangles(1))
    f = np.float32
    assert orc.unsigned_distance_f32(one, f([0.25]), f([0.25]), f([2]))[0] == 2.0
    assert orc.unsigned_distance_f32(one, f([2]), f([0]), f([0]))[0] == 1.0
    assert orc.unsigned_distance_f32(one, f([0.25]), f([0.25]), f([2]), culling_band=0.5)[0] == 0.5
    assert orc.unsigned_distance_f32(one, f([0.25]), f([0.25]), f([2]), culling_band=-4.0)[0] == 0.0
    twelve = orc.build_bvh4(synth.layered_triangles(12))
    d = orc.unsigned_distance_f32(twelve, f([0.25, 0.25, 0.0]), f([0.25, 0.25, 0.25]), f([10.25, 30, 10.25]),
                                  culling_band=2.0)
    assert d.tolist() == [0.25, 2.0, 0.25]
    four = orc.build_bvh4(synth.layered_triangles(4))
    assert four.internal_stack_bound == 0 and four.packed_stack_bound == 3
    assert orc.unsigned_distance_f32(four, f([0.25]), f([0.25]), f([3.5]), stack_capacity=1)[0] == 0.5


@pytest.mark.parametrize("name", ["torus", "soup"])
def test_distance_traversal_equals_brute_force(orc, name):
    if name == "torus":
        mesh = synth.torus(24, 12)
    else:
        rng = np.random.default_rng(8)
        v = rng.normal(size=(300, 3)).astype(np.float32)
        mesh = synth._mesh_from_arrays(v, rng.integers(0, 300, size=(200, 3)))
    bvh = orc.build_bvh4(mesh)
    qx, qy, qz = synth.uniform_box_points(4, 64, (-2.0,) * 3, (2.0,) * 3)
    got = orc.unsigned_distance_f32(bvh, qx, qy, qz)
    tris = bvh.triangles
    v0 = tris[:, 0:3]
    v1 = v0 + (tris[:, 3:6] - v0)  # the reference's (v0, e1, e2) reconstruction, in float32
    v2 = v0 + (tris[:, 6:9] - v0)
    for i in range(64):
        p = np.float32([qx[i], qy[i], qz[i]])
        best = min(orc.point_triangle_distance2_f32(p, v0[k], v1[k], v2[k]) for k in range(tris.shape[0]))
        # Culling only skips boxes no closer than the best so far: the traversal finds the same
        # minimum as scanning every triangle with the same float function.
        assert got[i] == np.float32(np.sqrt(np.float32(best)))


# ---------------------------------------------------------------------------------------------
# Ray first hit
# ---------------------------------------------------------------------------------------------

def _ray(orc, mesh, o, d, t_min=0.0, t_max=float("inf"), traversal="auto"):
    bvh = orc.build_bvh4(mesh)
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    return orc.ray_first_hit_f32(bvh, o[:, 0], o[:, 1], o[:, 2], d[:, 0], d[:, 1], d[:, 2], t_min, t_max,
                                 traversal=traversal)


def _mesh(x, y, z, i0, i1, i2):
    return synth._mesh_from_arrays(np.stack([x, y, z], 1), np.stack([i0, i1, i2], 1))


def test_ray_leaf_kats(orc):
    """tests/unit_gwn_query_distance_ray.cu:222-298."""
    one = synth.layered_triangles(1)
    h = _ray(orc, one, [0.25, 0.25, 1], [0, 0, -1])[0]
    assert h["status"] == 1 and h["primitive_id"] == 0 and h["t"] == 1.0
    assert abs(h["u"] - 0.25) <= 1e-6 and abs(h["v"] - 0.25) <= 1e-6
    assert (h["nx"], h["ny"], h["nz"]) == (0.0, 0.0, 1.0)
    assert _ray(orc, one, [0.25, 0.25, 1], [0, 0, -1], 1.0, 1.0)[0]["t"] == 1.0
    miss = _ray(orc, one, [0.25, 0.25, 1], [0, 0, 0], 0.0, 10.0)[0]
    assert miss["status"] == 0 and miss["t"] == -1.0 and miss["primitive_id"] == 0xFFFFFFFF
    h = _ray(orc, one, [0.25, 0.25, 1], [1e-30, 0, -1], 0.0, 2.0)[0]
    assert h["status"] == 1 and abs(h["t"] - 1) <= 1e-6
    assert _ray(orc, one, [0.25, 0.25, 0], [1, 0, 0], 0.0, 2.0)[0]["status"] == 0  # coplanar
    h = _ray(orc, one, [0.25, 0.25, 1e-6], [0, 0, -1], 0.0, 1e-5)[0]
    assert h["status"] == 1 and abs(h["t"] - 1e-6) <= 1e-10


@pytest.mark.parametrize("scale,tol", [(1e-10, 1e-6), (1e-20, 1e-5)])
def test_ray_tiny_and_subnormal_triangles(orc, scale, tol):
    """tests/unit_gwn_query_distance_ray.cu:300-348."""
    s = np.float32(scale)
    mesh = _mesh([0, s, 0], [0, 0, s], [0, 0, 0], [0], [1], [2])
    h = _ray(orc, mesh, [s * np.float32(0.25), s * np.float32(0.25), 1], [0, 0, -1], 0.0, 2.0)[0]
    assert h["status"] == 1
    assert abs(h["u"] - 0.25) <= tol and abs(h["v"] - 0.25) <= tol
    if scale == 1e-20:
        assert h["t"] == 1.0 and h["nz"] > 0


def test_ray_layered_and_octahedron_kats(orc):
    """tests/unit_gwn_query_distance_ray.cu:350-410, 465-482."""
    twelve = synth.layered_triangles(12)
    h = _ray(orc, twelve, [[0.25, 0.25, 20], [2, 2, 20], [0, 0.25, 20]], [[0, 0, -1]] * 3)
    assert h["t"].tolist() == [9.0, -1.0, 9.0]
    assert h["primitive_id"].tolist() == [11, 0xFFFFFFFF, 11]
    octa = _mesh([1, -1, 0, 0, 0, 0], [0, 0, 1, -1, 0, 0], [0, 0, 0, 0, 1, -1],
                 [0, 2, 1, 3, 2, 1, 3, 0], [2, 1, 3, 0, 0, 2, 1, 3], [4, 4, 4, 4, 5, 5, 5, 5])
    h = _ray(orc, octa, [[2, 2, 2], [2, 0.2, 0.1], [2, 2, 2]], [[-1, -1, -1], [-1, 0, 0], [1, 1, 1]])
    assert abs(h["t"][0] - 5 / 3) <= 1e-5 and h["primitive_id"][0] == 0
    assert abs(h["t"][1] - 1.3) <= 1e-5 and h["primitive_id"][1] == 0
    assert h["t"][2] == -1.0 and h["primitive_id"][2] == 0xFFFFFFFF
    sixteen = synth.layered_triangles(16)
    h = _ray(orc, sixteen, [0.25, 0.25, 20], [0, 0, 1], -30.0, 0.0)[0]
    assert h["status"] == 1 and h["primitive_id"] == 0 and h["t"] == -20.0


def test_ray_shared_edge_and_thin_bounds(orc):
    """tests/unit_gwn_query_distance_ray.cu:412-463."""
    quad = _mesh([0, 1, 1, 0], [0, 0, 1, 1], [0, 0, 0, 0], [0, 0], [1, 2], [2, 3])
    h = _ray(orc, quad, [0.5, 0.5, 1], [0, 0, -1], 0.0, 2.0)[0]
    assert h["status"] == 1 and h["t"] == 1.0 and h["primitive_id"] in (0, 1)
    assert np.isfinite(h["u"]) and np.isfinite(h["v"])
    thin = synth.layered_triangles(16)
    thin = synth.Mesh((thin.vx * np.float32(1e-25)).astype(np.float32), thin.vy, thin.vz, thin.i0, thin.i1, thin.i2)
    x = np.float32(0.25e-25)
    packed = _ray(orc, thin, [x, 0.25, -1], [0, 0, 1], traversal="packed")[0]
    canonical = _ray(orc, thin, [x, 0.25, -1], [0, 0, 1], traversal="canonical")[0]
    assert packed["status"] == 1 and packed["primitive_id"] == canonical["primitive_id"]
    assert packed["t"] == canonical["t"] == 1.0


def _moller_trumbore_f64(o, d, a, b, c):
    e1, e2 = b - a, c - a
    p = np.cross(d, e2)
    det = np.dot(e1, p)
    if abs(det) < 1e-300:
        return None
    inv = 1.0 / det
    s = o - a
    u = np.dot(s, p) * inv
    if u < 0 or u > 1:
        return None
    q = np.cross(s, e1)
    v = np.dot(d, q) * inv
    if v < 0 or u + v > 1:
        return None
    return np.dot(e2, q) * inv, u, v


def test_ray_matches_fp64_brute_force(orc):
    """tests/integration_gwn_static.cu:343-353: hit status and t against FP64 Moller-Trumbore
    (relative 5e-5) over every triangle; primitive id and barycentrics where the nearest hit is
    unambiguous."""
    mesh = synth.displaced_sphere(24, 12, seed=4)
    bvh = orc.build_bvh4(mesh)
    rng = np.random.default_rng(6)
    n = 256
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 3.0
    tgt = rng.uniform(-0.8, 0.8, size=(n, 3))
    d = tgt - o
    o32, d32 = o.astype(np.float32), d.astype(np.float32)
    hits = orc.ray_first_hit_f32(bvh, *(o32[:, k] for k in range(3)), *(d32[:, k] for k in range(3)))
    canonical = orc.ray_first_hit_f32(bvh, *(o32[:, k] for k in range(3)), *(d32[:, k] for k in range(3)),
                                      traversal="canonical")
    # Both traversals find the same nearest hit (no exact ties on this mesh).
    assert np.array_equal(hits["t"], canonical["t"]) and np.array_equal(hits["primitive_id"], canonical["primitive_id"])
    V = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float64)
    F = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.int64)
    for i in range(n):
        oo, dd = o32[i].astype(np.float64), d32[i].astype(np.float64)
        ts = []
        for f, (ia, ib, ic) in enumerate(F):
            r = _moller_trumbore_f64(oo, dd, V[ia], V[ib], V[ic])
            if r is not None and r[0] >= 0:
                ts.append((r[0], f, r[1], r[2]))
        ts.sort()
        assert hits["status"][i] == (1 if ts else 0)
        if not ts:
            continue
        assert abs(hits["t"][i] - ts[0][0]) <= 5e-5 * abs(ts[0][0])
        if len(ts) == 1 or ts[1][0] - ts[0][0] > 1e-4 * ts[0][0]:
            assert hits["primitive_id"][i] == ts[0][1]
            assert abs(hits["u"][i] - ts[0][2]) <= 1e-4 and abs(hits["v"][i] - ts[0][3]) <= 1e-4


# ---------------------------------------------------------------------------------------------
# Boundary chain
# ---------------------------------------------------------------------------------------------

def test_boundary_chain_kats(orc):
    """tests/unit_gwn_boundary.cu:94-187: empty, closed (cube) and open-cube chains, net
    orientation multiplicity 2 with the opposite direction cancelled."""
    s, e, m = orc.boundary_chain(synth._mesh_from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64)))
    assert s.size == 0
    s, e, m = orc.boundary_chain(synth.closed_cube())
    assert s.size == 0
    s, e, m = orc.boundary_chain(synth.open_cube())
    rows = sorted(zip(s.tolist(), e.tolist(), m.tolist()))
    assert rows == [(4, 7, 1), (5, 4, 1), (6, 5, 1), (7, 6, 1)]
    mesh = synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]),
                                   np.stack([[0, 1, 0, 0], [1, 0, 1, 1], [2, 3, 2, 2]], 1))
    s, e, m = orc.boundary_chain(mesh)
    rows = list(zip(s.tolist(), e.tolist(), m.tolist()))
    assert (0, 1, 2) in rows and not any(r[0] == 1 and r[1] == 0 for r in rows)


def test_boundary_chain_is_the_algebraic_boundary(orc):
    """Closed surfaces have no boundary; deleting faces leaves exactly the edges adjacent to one
    remaining face, oriented with it; rows come in (min vertex, max vertex) order."""
    assert orc.boundary_chain(synth.torus(16, 8))[0].size == 0
    assert orc.boundary_chain(synth.icosphere(2))[0].size == 0
    soup = synth.open_soup(16, 8, seed=2)
    s, e, m = orc.boundary_chain(soup)
    f = np.stack([soup.i0, soup.i1, soup.i2], 1).astype(np.int64)
    from collections import Counter
    net = Counter()
    for a, b, c in f:
        for u, v in ((a, b), (b, c), (c, a)):
            if u != v:
                net[(min(u, v), max(u, v))] += 1 if u < v else -1
    expect = sorted(((k[0], k[1], v) if v > 0 else (k[1], k[0], -v)) for k, v in net.items() if v != 0)
    assert sorted(zip(s.tolist(), e.tolist(), m.tolist())) == expect
    keys = [(min(a, b), max(a, b)) for a, b in zip(s.tolist(), e.tolist())]
    assert keys == sorted(keys)


# ---------------------------------------------------------------------------------------------
# Antipodal winding number
# ---------------------------------------------------------------------------------------------

def _antipodal(orc, mesh, qx, qy, qz, **kw):
    return orc.winding_antipodal_f32(orc.build_bvh4(mesh), mesh, orc.boundary_chain(mesh),
                                     np.float32(qx), np.float32(qy), np.float32(qz), **kw)


AP_Q = ([0.23, 2.2, -0.41, -2.3], [0.17, 0.3, 0.26, -0.4], [0.11, 0.2, 1.7, 0.5])


@pytest.mark.parametrize("meshname", ["closed_cube", "open_cube"])
def test_antipodal_matches_exact(orc, meshname):
    """tests/unit_gwn_query_antipodal.cu:332-396 (3e-4), including collinear boundary projections."""
    mesh = getattr(synth, meshname)()
    for q in (AP_Q, ([2, 2, -2, -2], [-1, 1, -1, 1], [2, 2, 2, 2])):
        got = _antipodal(orc, mesh, *q)
        ref = orc.winding_exact_mesh_f64(mesh, *(np.float32(v) for v in q))
        assert np.abs(got - ref).max() <= 3e-4


def test_antipodal_open_soup_matches_exact_both_precisions(orc):
    mesh = synth.open_soup(16, 8, seed=3)
    qx, qy, qz = synth.uniform_box_points(2, 256, (-1.4,) * 3, (1.4,) * 3)
    ref = orc.winding_exact_mesh_f64(mesh, qx, qy, qz)
    plain = _antipodal(orc, mesh, qx, qy, qz)
    split = _antipodal(orc, mesh, qx, qy, qz, near_fp64=True)
    assert np.abs(plain - ref).max() <= 3e-4
    assert np.abs(split - ref).max() <= 1e-6
    # Integer crossings + boundary term: the closed part contributes exactly 0 or 1.
    closed = _antipodal(orc, synth.icosphere(2), qx, qy, qz)
    exact_closed = orc.winding_exact_mesh_f64(synth.icosphere(2), qx, qy, qz)
    assert np.array_equal(closed, np.round(exact_closed).astype(np.float32))


def test_antipodal_gradient_matches_finite_differences(orc):
    """tests/unit_gwn_query_antipodal.cu:398-438 (6e-3) and against the exact FP64 gradient."""
    mesh = synth.open_cube()
    q = ([0.21, 1.3, -1.4, 0.35], [0.18, 0.25, -0.33, 1.4], [0.15, 0.2, 0.42, 1.55])
    w, gx, gy, gz = _antipodal(orc, mesh, *q, gradient=True)
    eps = 1e-2
    for axis, g in enumerate((gx, gy, gz)):
        p = [np.float32(v) for v in q]
        m = [np.float32(v) for v in q]
        p[axis] = p[axis] + np.float32(eps)
        m[axis] = m[axis] - np.float32(eps)
        fd = (_antipodal(orc, mesh, *p).astype(np.float64) - _antipodal(orc, mesh, *m)) / (2 * eps)
        assert np.abs(g - fd).max() <= 6e-3
    ref = orc.winding_gradient_exact_mesh_f64(mesh, *(np.float32(v) for v in q))
    for g, r in zip((gx, gy, gz), ref):
        assert np.abs(g - r).max() <= 1e-5


def test_antipodal_gradient_axis_retry_symmetry(orc):
    """tests/unit_gwn_query_antipodal.cu:440-460: queries on the x = 1 face plane project a
    boundary edge through the query for Z; the retried axis gives mirror-symmetric gradients."""
    mesh = synth.open_cube()
    q = ([1, 1, 1, 1], [0.25, -0.25, 0.5, -0.5], [0.2, 0.2, 0.4, 0.4])
    _, gx, gy, gz = _antipodal(orc, mesh, *q, gradient=True)
    assert all(np.all(np.isfinite(g)) for g in (gx, gy, gz))
    for a, b in ((0, 1), (2, 3)):
        assert abs(gx[a] - gx[b]) <= 1e-6 and abs(gy[a] + gy[b]) <= 1e-6 and abs(gz[a] - gz[b]) <= 1e-6


def test_antipodal_all_axes_singular_is_nan(orc):
    """A query on a boundary vertex projects an edge through itself on every axis."""
    mesh = synth.open_cube()
    w = _antipodal(orc, mesh, [1.0], [1.0], [1.0])
    assert np.isnan(w[0])


# ---------------------------------------------------------------------------
# H-PLOC binary builder (include/gwn/detail/gwn_bvh_build_hploc.cuh:104-622)
# ---------------------------------------------------------------------------

def _equal_triangles(n):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float32)
    return synth._mesh_from_arrays(v, np.tile([[0, 1, 2]], (n, 1)))


def _children(bvh, node):
    l, r, lk, rk = (int(x) for x in bvh.binary_nodes[node])
    return (l, lk), (r, rk)


def _leaf_sets(bvh):
    """Sorted-leaf set of every binary internal node, from the node arrays alone."""
    T = bvh.binary_nodes.shape[0] + 1
    memo = {}

    def rec(node):
        if node in memo:
            return memo[node]
        out = []
        for idx, kind in _children(bvh, node):
            out.extend([idx] if kind == 2 else rec(idx))
        memo[node] = sorted(out)
        return memo[node]

    import sys

    sys.setrecursionlimit(max(10000, 4 * T))
    rec(bvh.binary_root)
    return memo


def _check_complete_binary_tree(bvh, T):
    """2T - 1 clusters: every leaf and every non-root internal node is a child exactly once
    (hploc.cuh:613-619), bounds are unions of the children, first leaf is the smallest leaf."""
    leaf_seen = np.zeros(T, np.int64)
    node_seen = np.zeros(T - 1, np.int64)
    for n in range(T - 1):
        for idx, kind in _children(bvh, n):
            assert kind in (1, 2)
            (leaf_seen if kind == 2 else node_seen)[idx] += 1
    assert np.all(leaf_seen == 1)
    expect_nodes = np.ones(T - 1, np.int64)
    expect_nodes[bvh.binary_root] = 0
    assert np.array_equal(node_seen, expect_nodes)
    sets = _leaf_sets(bvh)
    assert len(sets) == T - 1 and sorted(sets[bvh.binary_root]) == list(range(T))
    rec = np.asarray(bvh.triangles, np.float32)
    lo = np.minimum(np.minimum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
    hi = np.maximum(np.maximum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
    for n, leaves in sets.items():
        assert int(bvh.binary_first[n]) == leaves[0]
        box = np.concatenate([lo[leaves].min(axis=0), hi[leaves].max(axis=0)])
        assert np.array_equal(np.asarray(bvh.binary_bounds[n], np.float32).view(np.uint32), box.view(np.uint32))


@pytest.mark.parametrize("radius", [1, 4, 8])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 9])
def test_hploc_equal_boxes_build_the_perfect_tree(orc, k, radius):
    """Identical boxes make every union area equal, so the packed key's offset and direction
    parity decide (hploc.cuh:195-241): even lanes take their right neighbour, odd lanes their left
    one, and a round pairs (0,1), (2,3), ...  Equal Morton codes fall back to positions, whose XOR
    deltas make the implicit hierarchy the perfect tree on aligned blocks when T = 2^k.  The whole
    tree is then the perfect tree: the node over [a, a + 2s) has children [a, a + s) and
    [a + s, a + 2s) (lower lane on the left) and, numbered by its right child's smallest leaf,
    id a + s - 1."""
    T = 1 << k
    bvh = orc.build_bvh4(_equal_triangles(T), method="hploc", radius=radius)
    assert bvh.binary_root == T // 2 - 1
    s = 1
    while s < T:
        for a in range(0, T, 2 * s):
            left = (a, 2) if s == 1 else (a + s // 2 - 1, 1)
            right = (a + s, 2) if s == 1 else (a + s + s // 2 - 1, 1)
            assert _children(bvh, a + s - 1) == (left, right), (T, a, s)
        s *= 2


def _half_area32(lo, hi):
    d = np.maximum(hi - lo, np.float32(0))
    return np.float32(d[0] * d[1]) + np.float32(d[1] * d[2]) + np.float32(d[2] * d[0])


def _ploc_by_definition(lo, hi, radius):
    """PLOC (Meister and Bittner 2018) on clusters in sorted order: each cluster picks the
    neighbour within `radius` positions whose union box has the smallest half-area; mutual pairs
    merge into the lower position; survivors keep their order; rounds repeat until one cluster
    remains.  Returns the merges as (left leaf set, right leaf set) and whether two candidate
    areas came within 2^-19 of each other (where the device key's dropped low bits could decide)."""
    clusters = [([i], lo[i].copy(), hi[i].copy()) for i in range(lo.shape[0])]
    merges, near_tie = [], False
    while len(clusters) > 1:
        n = len(clusters)
        nn = []
        for i in range(n):
            cand = []
            for j in range(max(0, i - radius), min(n, i + radius + 1)):
                if j != i:
                    a = _half_area32(np.minimum(clusters[i][1], clusters[j][1]),
                                     np.maximum(clusters[i][2], clusters[j][2]))
                    cand.append((float(a), j))
            cand.sort()
            if len(cand) > 1 and cand[1][0] - cand[0][0] <= cand[0][0] * 2.0 ** -19:
                near_tie = True
            nn.append(cand[0][1])
        out = []
        for i in range(n):
            j = nn[i]
            if nn[j] == i:
                if i < j:
                    merges.append((clusters[i][0], clusters[j][0]))
                    out.append((clusters[i][0] + clusters[j][0], np.minimum(clusters[i][1], clusters[j][1]),
                                np.maximum(clusters[i][2], clusters[j][2])))
            else:
                out.append(clusters[i])
        clusters = out
    return merges, near_tie


@pytest.mark.parametrize("radius", [1, 2, 4, 8])
@pytest.mark.parametrize("T", [2, 3, 7, 12, 16])
def test_hploc_small_inputs_equal_ploc_by_definition(orc, T, radius):
    """Up to 16 primitives no partial range reaches the merge threshold, so H-PLOC is PLOC on the
    whole sorted sequence (hploc.cuh:420-447, 451-476)."""
    checked = 0
    for seed in range(6):
        rng = np.random.default_rng(100 + 17 * T + seed)
        v = rng.uniform(-1, 1, size=(3 * T, 3)).astype(np.float32)
        mesh = synth._mesh_from_arrays(v, np.arange(3 * T).reshape(T, 3))
        bvh = orc.build_bvh4(mesh, method="hploc", radius=radius)
        rec = np.asarray(bvh.triangles, np.float32)
        lo = np.minimum(np.minimum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
        hi = np.maximum(np.maximum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
        merges, near_tie = _ploc_by_definition(lo, hi, radius)
        if near_tie:
            continue
        sets = _leaf_sets(bvh)
        got = set()
        for n, leaves in sets.items():
            (l, lk), (r, rk) = _children(bvh, n)
            left = [l] if lk == 2 else sets[l]
            right = [r] if rk == 2 else sets[r]
            got.add((tuple(sorted(left)), tuple(sorted(right))))
        assert got == {(tuple(sorted(a)), tuple(sorted(b))) for a, b in merges}
        checked += 1
    assert checked >= 3


def _sah(bvh, T):
    areas = [float(_half_area32(np.asarray(b[:3], np.float32), np.asarray(b[3:], np.float32)))
             for b in bvh.binary_bounds]
    return sum(areas) / areas[bvh.binary_root]


@pytest.mark.parametrize("radius", [1, 4, 8])
@pytest.mark.parametrize("name", ["displaced", "soup", "torus", "equal513"])
def test_hploc_tree_is_complete_and_canonical(orc, name, radius):
    mesh = {
        "displaced": lambda: synth.displaced_sphere(48, 24),
        "soup": lambda: synth.open_soup(40, 20, seed=7),
        "torus": lambda: synth.torus(32, 16),
        "equal513": lambda: _equal_triangles(513),
    }[name]()
    T = mesh.i0.shape[0]
    bvh = orc.build_bvh4(mesh, method="hploc", radius=radius)
    _check_complete_binary_tree(bvh, T)
    helpers.check_bvh4(bvh.nodes, bvh.root_reference, bvh.primitive_indices, bvh.triangles, mesh,
                       bvh.internal_stack_bound, bvh.packed_stack_bound)
    helpers.check_collapse_rule(bvh.nodes, bvh.binary_nodes, bvh.binary_bounds, bvh.binary_root)
    lb = orc.build_bvh4(mesh)
    assert np.array_equal(bvh.primitive_indices, lb.primitive_indices)
    assert np.array_equal(bvh.triangles, lb.triangles)
    if name in ("displaced", "torus") and radius >= 4:
        # Locally ordered clustering yields a lower surface-area cost than the Morton split tree.
        assert _sah(bvh, T) < _sah(lb, T)


def test_wide_numbering_is_breadth_first_by_first_leaf(orc):
    """The canonical order: breadth-first, parents in level order, siblings by smallest sorted
    leaf.  For LBVH, whose subtrees are contiguous leaf ranges, that is (depth, first leaf)."""
    for method in ("lbvh", "hploc"):
        mesh = synth.displaced_sphere(48, 24)
        bvh = orc.build_bvh4(mesh, method=method)
        nodes = helpers.as_nodes(bvh.nodes)
        rec = np.asarray(bvh.triangles)
        first = {}
        depth = {0: 0}

        def first_leaf(n):
            if n in first:
                return first[n]
            best = None
            for ch in nodes[n]["children"]:
                ref = int(ch["reference"])
                if not ref & helpers.VALID_BIT:
                    continue
                f = (ref & helpers.OFFSET_MASK) if (ref >> 48) else first_leaf(ref & helpers.OFFSET_MASK)
                best = f if best is None else min(best, f)
            first[n] = best
            return best

        expect = []
        queue = [0]
        while queue:
            n = queue.pop(0)
            expect.append(n)
            kids = [int(ch["reference"]) & helpers.OFFSET_MASK for ch in nodes[n]["children"]
                    if int(ch["reference"]) & helpers.VALID_BIT and (int(ch["reference"]) >> 48) == 0]
            queue.extend(sorted(kids, key=first_leaf))
        assert expect == list(range(len(nodes)))
        if method == "lbvh":
            for n in range(len(nodes)):
                for ch in nodes[n]["children"]:
                    ref = int(ch["reference"])
                    if ref & helpers.VALID_BIT and (ref >> 48) == 0:
                        depth[ref & helpers.OFFSET_MASK] = depth[n] + 1
            keys = [(depth[n], first_leaf(n)) for n in range(len(nodes))]
            assert keys == sorted(keys)


@pytest.mark.parametrize("count", [15, 16, 17, 31, 32, 33, 63, 64, 65, 129, 513])
def test_hploc_repeated_equal_bounds_winding(orc, count):
    """tests/integration_gwn_builder.cu:301-343: identical triangles at counts around the list
    threshold; every radius builds a complete tree whose fully descended Taylor winding equals the
    exact winding (2e-5 + 2e-5 |w|)."""
    mesh = _equal_triangles(count)
    q = (np.float32([0.2, 0.7, -0.3]), np.float32([0.2, 0.1, 0.3]), np.float32([1, -2, 0.5]))
    ref = orc.winding_exact_mesh_f64(mesh, *q)
    for radius in (1, 4, 8):
        bvh = orc.build_bvh4(mesh, method="hploc", radius=radius)
        _check_complete_binary_tree(bvh, count)
        w = orc.winding_taylor_f32(bvh, orc.refit_moment(bvh, 2), 2, *q, accuracy_scale=1e6)
        assert np.all(np.abs(w - ref) <= 2e-5 + 2e-5 * np.abs(ref))


def test_hploc_rejects_bad_radius(orc):
    mesh = _equal_triangles(4)
    for radius in (0, 9):
        with pytest.raises(RuntimeError):
            orc.build_bvh4(mesh, method="hploc", radius=radius)


# ---------------------------------------------------------------------------
# Adjoint of winding-number and distance losses (FP64 oracle)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("beta", [2.0, 1e6])
@pytest.mark.parametrize("order", [0, 1])
def test_taylor_adjoint_matches_central_differences(orc, order, beta):
    """dL/dV of L = sum_q u_q w_q against FP64 central differences of the same Taylor sum (same
    accept/descend decisions) on 64 random vertex coordinates; beta = 1e6 descends every leaf, so
    that case checks the exact solid-angle derivative alone.  Queries close to the surface make the
    difference quotient converge slowly (at h = 1e-4 one coordinate is still 5e-5 off), so the
    step is 1e-6 with Richardson extrapolation; FP64 roundoff there is about 1e-8."""
    mesh = synth.open_soup(12, 6, seed=3)
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, order)
    rng = np.random.default_rng(order)
    qx, qy, qz = synth.uniform_box_points(5, 300, (-1.4,) * 3, (1.4,) * 3)
    u = rng.normal(size=300)
    dv, _ = orc.winding_taylor_adjoint_f64(bvh, mom, order, beta, mesh, qx, qy, qz, u)
    V0 = np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)

    def central(k, vi, h):
        Vp, Vm = V0.copy(), V0.copy()
        Vp[k, vi] += h
        Vm[k, vi] -= h
        return (orc.winding_taylor_frozen_f64(bvh, mom, order, beta, mesh, qx, qy, qz, u, Vp) -
                orc.winding_taylor_frozen_f64(bvh, mom, order, beta, mesh, qx, qy, qz, u, Vm)) / (2 * h)

    h = 1e-6
    for flat in rng.choice(V0.size, 64, replace=False):
        k, vi = divmod(int(flat), V0.shape[1])
        # Richardson extrapolation of the h and h/2 central differences cancels the h^2 term.
        fd = (4 * central(k, vi, h / 2) - central(k, vi, h)) / 3
        assert abs(fd - dv[k, vi]) <= 1e-6 * (1 + abs(fd)), (k, vi, fd, dv[k, vi])


@pytest.mark.parametrize("order", [0, 1])
def test_taylor_adjoint_query_gradient_is_the_winding_gradient(orc, order):
    """dL/dq = u_q * grad_q w_q: the adjoint's query part equals the Taylor gradient replay."""
    mesh = synth.open_soup(12, 6, seed=3)
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, order)
    qx, qy, qz = synth.uniform_box_points(6, 400, (-1.4,) * 3, (1.4,) * 3)
    u = np.random.default_rng(7).normal(size=400)
    _, dq = orc.winding_taylor_adjoint_f64(bvh, mom, order, 2.0, mesh, qx, qy, qz, u)
    _, gx, gy, gz = orc.winding_gradient_taylor_f32(bvh, mom, order, qx, qy, qz, accuracy_scale=2.0)
    g = np.stack([gx, gy, gz], 1).astype(np.float64) * u[:, None]
    assert np.abs(dq - g).max() <= 1e-5 * (1 + np.abs(g).max())


def test_taylor_adjoint_translation_and_closed_mesh_invariants(orc):
    """Moving the mesh and the queries together leaves w unchanged: sum of dL/dV + sum of dL/dq = 0
    per axis; and the exact winding of a closed mesh is locally constant away from its surface, so
    with every leaf descended dL/dV vanishes for queries far from it."""
    mesh = synth.open_soup(12, 6, seed=4)
    bvh = orc.build_bvh4(mesh)
    mom = orc.refit_moment(bvh, 1)
    qx, qy, qz = synth.uniform_box_points(8, 300, (-1.4,) * 3, (1.4,) * 3)
    u = np.random.default_rng(2).normal(size=300)
    dv, dq = orc.winding_taylor_adjoint_f64(bvh, mom, 1, 2.0, mesh, qx, qy, qz, u)
    assert np.abs(dv.sum(axis=1) + dq.sum(axis=0)).max() <= 1e-9 * (1 + np.abs(dv).sum())
    closed = synth.icosphere(2)
    cb = orc.build_bvh4(closed)
    far = synth.uniform_box_points(9, 200, (2.0,) * 3, (3.0,) * 3)
    dv2, _ = orc.winding_taylor_adjoint_f64(cb, orc.refit_moment(cb, 1), 1, 1e6, closed, *far, np.ones(200))
    assert np.abs(dv2).max() <= 1e-10


def test_distance_adjoint_matches_central_differences(orc):
    """L = sum_p u_p d_p: dL/dV against FP64 central differences of brute-force distances, dL/dp
    is u_p times a unit vector, and the distances equal the traversal oracle's."""
    mesh = synth.open_soup(12, 6, seed=3)
    rng = np.random.default_rng(1)
    px, py, pz = synth.uniform_box_points(7, 400, (-1.4,) * 3, (1.4,) * 3)
    u = rng.normal(size=400)
    dist, dv, dp = orc.distance_adjoint_f64(mesh, px, py, pz, u)
    V0 = np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)
    h = 1e-6
    for flat in rng.choice(V0.size, 64, replace=False):
        k, vi = divmod(int(flat), V0.shape[1])
        Vp, Vm = V0.copy(), V0.copy()
        Vp[k, vi] += h
        Vm[k, vi] -= h
        fd = (np.dot(u, orc.distance_adjoint_f64(mesh, px, py, pz, u, vertices=Vp)[0]) -
              np.dot(u, orc.distance_adjoint_f64(mesh, px, py, pz, u, vertices=Vm)[0])) / (2 * h)
        assert abs(fd - dv[k, vi]) <= 1e-6 * (1 + abs(fd))
    assert np.allclose(np.linalg.norm(dp / u[:, None], axis=1), 1.0, atol=1e-12)
    ref = orc.unsigned_distance_f32(orc.build_bvh4(mesh), px, py, pz)
    assert np.abs(dist - ref).max() <= 1e-6


def test_adjoint_magnitude_single_terms_and_additivity(orc):
    """The envelope sums |u| |dW_t/dv| and |u| |dW_t/dq|: for one triangle and one query they are the
    absolute central differences of the FP64 solid angle and the absolute exact gradient; over a
    mesh they add triangle by triangle and scale with |u|."""
    tri = np.float64([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.2, 0.9, 0.1]])
    one = synth._mesh_from_arrays(tri.astype(np.float32), [[0, 1, 2]])
    V = np.stack([one.vx, one.vy, one.vz]).astype(np.float64)
    q = np.float64([0.3, 0.2, 0.4])
    u = np.float64([-1.5])
    mv, mq = orc.adjoint_magnitude_f64(one, *(np.float32([c]) for c in q), u)
    qf = q.astype(np.float32).astype(np.float64)
    h = 1e-6
    for vi in range(3):
        for k in range(3):
            Vp, Vm = V.copy(), V.copy()
            Vp[k, vi] += h
            Vm[k, vi] -= h
            fd = (orc.solid_angle_f64(*Vp.T, qf) - orc.solid_angle_f64(*Vm.T, qf)) / (2 * h) / (4 * math.pi)
            assert abs(mv[k, vi] - 1.5 * abs(fd)) <= 1e-7 * (1 + abs(fd))
    g = np.array(orc.winding_gradient_exact_mesh_f64(one, *(np.float32([c]) for c in q))).ravel()
    assert np.allclose(mq[0], 1.5 * np.abs(g), rtol=1e-12, atol=0)
    # Two disjoint copies: per-vertex sums equal the single-triangle sums, per-query sums add.
    two = synth._mesh_from_arrays(np.concatenate([tri, tri + [0.0, 0.0, 2.0]]).astype(np.float32),
                                  [[0, 1, 2], [3, 4, 5]])
    q2 = synth.uniform_box_points(3, 16, (-1.0,) * 3, (1.0,) * 3)
    u2 = np.random.default_rng(2).normal(size=16)
    mv2, mq2 = orc.adjoint_magnitude_f64(two, *q2, u2)
    parts = [synth._mesh_from_arrays((tri + [0.0, 0.0, 2.0 * s]).astype(np.float32), [[0, 1, 2]]) for s in (0, 1)]
    m = [orc.adjoint_magnitude_f64(p, *q2, u2) for p in parts]
    assert np.allclose(mv2, np.concatenate([m[0][0], m[1][0]], 1), rtol=1e-12, atol=0)
    assert np.allclose(mq2, m[0][1] + m[1][1], rtol=1e-12, atol=0)
    assert np.allclose(orc.adjoint_magnitude_f64(two, *q2, -3 * u2)[0], 3 * mv2, rtol=1e-12, atol=0)


# ---- widths 2 and 3 (gwn_bvh_object<Width>, tests/integration_gwn_builder.cu:256-298) ----------

@pytest.mark.parametrize("method", ["lbvh", "hploc"])
def test_width_two_tree_is_the_binary_tree(orc, method):
    """Width 2 opens no binary node: every wide node is one binary node with its two children,
    so the node count is T - 1 and the children are the binary children in (left, right) order."""
    mesh = synth.torus(16, 9)
    bvh = orc.build_bvh4(mesh, method, 4, width=2)
    T = mesh.triangle_count
    assert bvh.node_count == T - 1
    helpers.check_bvh4(bvh.nodes, bvh.root_reference, bvh.primitive_indices, bvh.triangles, mesh,
                       bvh.internal_stack_bound, bvh.packed_stack_bound)
    helpers.check_collapse_rule(bvh.nodes, bvh.binary_nodes, bvh.binary_bounds, bvh.binary_root, width=2)
    nodes = helpers.as_nodes(bvh.nodes)
    for n in range(bvh.node_count):
        refs = [int(c["reference"]) for c in nodes[n]["children"]]
        assert all(r & helpers.VALID_BIT for r in refs[:2]) and not any(r & helpers.VALID_BIT for r in refs[2:])
    # A binary tree's stack bounds: one deferred child per level.
    assert bvh.packed_stack_bound >= bvh.internal_stack_bound


@pytest.mark.parametrize("width", [2, 3])
@pytest.mark.parametrize("method", ["lbvh", "hploc"])
def test_narrow_trees_keep_leaves_and_expansion_rule(orc, width, method):
    full = synth.displaced_sphere(24, 12)
    half = full.triangle_count // 2  # an open cap: its summed normal is far from zero
    mesh = synth._mesh_from_arrays(np.stack([full.vx, full.vy, full.vz], 1),
                                   np.stack([full.i0, full.i1, full.i2], 1)[:half])
    wide = orc.build_bvh4(mesh, method, 8, width=4)
    bvh = orc.build_bvh4(mesh, method, 8, width=width)
    # Same sort and binary hierarchy; only the collapse differs.
    assert np.array_equal(bvh.primitive_indices, wide.primitive_indices)
    assert np.array_equal(bvh.binary_nodes, wide.binary_nodes)
    helpers.check_bvh4(bvh.nodes, bvh.root_reference, bvh.primitive_indices, bvh.triangles, mesh,
                       bvh.internal_stack_bound, bvh.packed_stack_bound)
    helpers.check_collapse_rule(bvh.nodes, bvh.binary_nodes, bvh.binary_bounds, bvh.binary_root, width=width)
    # Every node of a width-w tree opens w - 1 binary nodes unless its frontier is all leaves.
    assert bvh.node_count > wide.node_count
    # The summed area-weighted normal N (fields 4-6, additive) over the root's slots is the whole
    # mesh's in every width (gwn_bvh.cuh:144-152).
    for order in (0, 1, 2):
        m = orc.refit_moment(bvh, order, dtype=np.float64)
        mw = orc.refit_moment(wide, order, dtype=np.float64)
        n_root, n_wide = m[0, 4:7].sum(axis=1), mw[0, 4:7].sum(axis=1)
        assert np.linalg.norm(n_wide) > 1.0
        assert np.allclose(n_root, n_wide, rtol=1e-12, atol=1e-12)


# ---- H-PLOC on generic inputs, restated from its definition -------------------------------------

def _hploc_by_definition(codes, lo, hi, radius, threshold=16):
    """H-PLOC (Benthin et al. 2024, hploc.cuh:104-478) restated top-down: the implicit hierarchy is
    the binary radix tree of the sorted (code, position) keys (helpers.lbvh_topdown, independent of
    the oracle's bottom-up climb over range boundaries); every subtree, children first, gets the
    cluster list of its children's lists concatenated (a leaf is a one-cluster list), and a subtree
    of more than `threshold` leaves, or the whole tree, runs PLOC on it until at most `threshold`
    clusters (one for the whole tree) remain.  PLOC's neighbour key is the device's packed key
    (hploc.cuh:195-241): union half-area bits << 1 with the low four bits replaced by
    (offset - 1) << 1 and the direction parity; mutual pairs merge into the lower position.
    Returns the merges as ordered (left leaf set, right leaf set)."""
    T = len(codes)
    tree = helpers.lbvh_topdown(codes)

    def key(a, b):
        area = helpers.half_area32(np.concatenate([np.minimum(a[1], b[1]), np.maximum(a[2], b[2])]))
        return (int(np.float32(area).view(np.uint32)) << 1) & 0xFFFFFFF0

    merges = []

    def ploc(lst, final):
        limit = 1 if final else threshold
        while len(lst) > limit:
            n = len(lst)
            nn = []
            for l in range(n):
                best = None
                for r in range(1, radius + 1):
                    if l + r < n:
                        k = key(lst[l], lst[l + r]) | ((r - 1) << 1) | (l & 1)
                        best = k if best is None else min(best, k)
                    if l >= r:
                        k = key(lst[l - r], lst[l]) | ((r - 1) << 1) | ((l & 1) ^ 1)
                        best = k if best is None else min(best, k)
                off = ((best & 15) >> 1) + 1
                nn.append(l + off if ((best ^ l) & 1) == 0 else l - off)
            out = []
            for l in range(n):
                j = nn[l]
                if nn[j] == l:
                    if l < j:
                        a, b = lst[l], lst[j]
                        merges.append((tuple(a[0]), tuple(b[0])))
                        out.append((sorted(a[0] + b[0]), np.minimum(a[1], b[1]), np.maximum(a[2], b[2])))
                else:
                    out.append(lst[l])
            lst = out
        return lst

    def leaf(i):
        return ([i], lo[i].astype(np.float32), hi[i].astype(np.float32))

    lists, sizes = {}, {}
    order = []
    stack = [0]
    while stack:  # post-order: children before parents
        n = stack.pop()
        order.append(n)
        left, left_leaf, right, right_leaf = tree[n]
        if not left_leaf:
            stack.append(left)
        if not right_leaf:
            stack.append(right)
    for n in reversed(order):
        left, left_leaf, right, right_leaf = tree[n]
        ll = [leaf(left)] if left_leaf else lists[left]
        rl = [leaf(right)] if right_leaf else lists[right]
        size = (1 if left_leaf else sizes[left]) + (1 if right_leaf else sizes[right])
        lst = ll + rl
        if size > threshold or n == 0:
            lst = ploc(lst, final=n == 0)
        lists[n], sizes[n] = lst, size
    assert sizes[0] == T and len(lists[0]) == 1
    return merges


@pytest.mark.parametrize("radius", [1, 4, 8])
@pytest.mark.parametrize("name", ["displaced", "soup", "torus"])
def test_hploc_merges_equal_the_restated_definition(orc, name, radius):
    """Every merge of the oracle's H-PLOC on real meshes (thousands of triangles, so ranges above the
    16-cluster threshold run PLOC repeatedly) equals the top-down restatement's, as ordered
    (lower-position, higher-position) cluster pairs, and each internal node's id is the larger of
    its two clusters' smallest sorted leaves minus one."""
    mesh = {
        "displaced": lambda: synth.displaced_sphere(40, 20),
        "soup": lambda: synth.open_soup(36, 18, seed=7),
        "torus": lambda: synth.torus(28, 14),
    }[name]()
    bvh = orc.build_bvh4(mesh, method="hploc", radius=radius)
    T = mesh.triangle_count
    rec = np.asarray(bvh.triangles, np.float32)
    lo = np.minimum(np.minimum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
    hi = np.maximum(np.maximum(rec[:, 0:3], rec[:, 3:6]), rec[:, 6:9])
    want = _hploc_by_definition(bvh.sorted_morton, lo, hi, radius)
    assert len(want) == T - 1
    sets = _leaf_sets(bvh)
    got = {}
    for n in sets:
        (l, lk), (r, rk) = _children(bvh, n)
        left = (l,) if lk == 2 else tuple(sets[l])
        right = (r,) if rk == 2 else tuple(sets[r])
        got[(left, right)] = n
        assert n == max(left[0], right[0]) - 1
    assert set(got) == {(tuple(sorted(a)), tuple(sorted(b))) for a, b in want}


# ---------------------------------------------------------------------------------------------
# Query presort key (test-side restatement of the Hilbert curve)
# ---------------------------------------------------------------------------------------------

@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_hilbert_key_visits_every_cell_by_face_steps(bits):
    """On a 2^b grid (the 1024^3 grid's top b bits) the keys are a bijection onto [0, 8^b) and
    consecutive keys are face-adjacent cells: the defining property of the Hilbert curve, which
    neither a Morton interleave nor a transposed or mis-ordered Gray step satisfies."""
    n = 1 << bits
    g = np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij")).reshape(3, -1)
    shift = 10 - bits
    key = helpers.hilbert_key10(*(g << shift)) >> np.uint64(3 * shift)
    assert np.array_equal(np.sort(key), np.arange(n ** 3, dtype=np.uint64))
    walk = g[:, np.argsort(key)]
    assert np.all(np.abs(np.diff(walk, axis=1)).sum(axis=0) == 1)
    assert tuple(walk[:, 0]) == (0, 0, 0)


def test_hilbert_key_blocks_at_full_resolution():
    """At 10 bits, every aligned 2^L block (L = 1, 2, 3) at random places holds a contiguous run
    of 8^L keys walked by face steps: the curve's recursion is right down to the lowest bits."""
    rng = np.random.default_rng(3)
    for level in (1, 2, 3):
        m = 1 << level
        local = np.stack(np.meshgrid(*(np.arange(m),) * 3, indexing="ij")).reshape(3, -1)
        for _ in range(16):
            origin = rng.integers(0, 1024 // m, size=3)[:, None] * m
            cells = local + origin
            key = helpers.hilbert_key10(*cells).astype(np.int64)
            ks = np.sort(key)
            assert np.array_equal(ks, np.arange(ks[0], ks[0] + m ** 3))
            walk = cells[:, np.argsort(key)]
            assert np.all(np.abs(np.diff(walk, axis=1)).sum(axis=0) == 1)
