"""Pins the oracle's Real = double functions (exact sum, double tree, double moments, double Taylor)
before they check the HIP library's f64 entry points.

Each check comes from outside the functions' own formulas: closed forms (the octant triangle's
solid angle, a closed mesh's winding number), the float-record oracle on float-representable
inputs (the same arithmetic on the same values), an exhaustive min/max over each subtree, and the
limit in which the Taylor traversal descends everywhere and must equal the exact sum.
"""

import numpy as np
import pytest

from smallgwn_amd import synth


def _vertices(mesh):
    return np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)


def _perturbed(mesh, seed=11, rel=1e-9):
    """Double vertices that are not float values: each coordinate moved by ~rel of its size."""
    v = _vertices(mesh)
    u = synth.uniform01(seed, v.size).reshape(v.shape) * 2.0 - 1.0
    return v * (1.0 + rel * u)


def test_exact_d64_closed_forms(orc):
    # Octant triangle seen from the origin: solid angle pi/2, so |w| = 1/8
    # (the spherical triangle of the first octant).
    v = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    w = orc.winding_exact_mesh_d64(v, [0], [1], [2], [0.0], [0.0], [0.0])
    assert abs(abs(w[0]) - 0.125) <= 1e-15
    # A closed, outward-oriented mesh: 1 inside, 0 outside, in double, for double vertices.
    mesh = synth.icosphere(3)
    vd = _perturbed(mesh)
    q = (np.array([0.0, 0.1, 3.0, -2.5]), np.array([0.0, -0.2, 0.0, 1.0]), np.array([0.0, 0.3, 0.0, 0.5]))
    w = orc.winding_exact_mesh_d64(vd, mesh.i0, mesh.i1, mesh.i2, *q)
    assert np.all(np.abs(w - np.array([1.0, 1.0, 0.0, 0.0])) <= 1e-12)


def test_exact_d64_equals_float_vertex_oracle_on_float_values(orc):
    mesh = synth.torus(24, 12)
    q = synth.uniform_box_points(5, 256, (-1.5,) * 3, (1.5,) * 3)
    a = orc.winding_exact_mesh_d64(_vertices(mesh), mesh.i0, mesh.i1, mesh.i2, *q)
    b = orc.winding_exact_mesh_f64(mesh, *q)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("width", [2, 4])
def test_double_tree_boxes_are_subtree_unions(orc, width):
    mesh = synth.displaced_sphere(32, 16)
    bvh = orc.build_bvh4(mesh, "hploc", 3, width=width)
    vd = _perturbed(mesh)
    rec, child, root = orc.double_tree(bvh, vd, mesh.i0, mesh.i1, mesh.i2)
    # Records are the double vertices of each sorted primitive.
    f = np.stack([mesh.i0, mesh.i1, mesh.i2]).astype(np.int64)[:, bvh.primitive_indices]
    assert np.array_equal(rec, np.concatenate([vd[:, f[0]].T, vd[:, f[1]].T, vd[:, f[2]].T], axis=1))
    corners = rec.reshape(-1, 3, 3)
    lo_t, hi_t = corners.min(axis=1), corners.max(axis=1)

    def leaves(n):  # exhaustive: every leaf offset under node n
        out = []
        for s in range(4):
            ref = int(bvh.nodes[n]["children"][s]["reference"])
            if not ref >> 47 & 1:
                continue
            off = ref & ((1 << 47) - 1)
            out += leaves(off) if ref >> 48 == 0 else list(range(off, off + (ref >> 48)))
        return out

    for n in range(bvh.node_count):
        for s in range(4):
            ref = int(bvh.nodes[n]["children"][s]["reference"])
            if not ref >> 47 & 1:
                assert not child[n, s].any()
                continue
            off = ref & ((1 << 47) - 1)
            ids = leaves(off) if ref >> 48 == 0 else list(range(off, off + (ref >> 48)))
            assert np.array_equal(child[n, s, :3], lo_t[ids].min(axis=0))
            assert np.array_equal(child[n, s, 3:], hi_t[ids].max(axis=0))
    assert np.array_equal(root, np.concatenate([vd.min(axis=1), vd.max(axis=1)]))
    # On float values the double boxes are the float build's boxes.
    _, child32, _ = orc.double_tree(bvh, _vertices(mesh), mesh.i0, mesh.i1, mesh.i2)
    b32 = np.stack([[bvh.nodes[n]["children"][s]["bounds"] for s in range(4)] for n in range(bvh.node_count)])
    assert np.array_equal(child32, b32.astype(np.float64))


@pytest.mark.parametrize("order", [0, 1, 2])
def test_moment_d64_equals_float_record_oracle_on_float_values(orc, order):
    mesh = synth.displaced_sphere(32, 16)
    bvh = orc.build_bvh4(mesh, "hploc", 8)
    rec, child, _ = orc.double_tree(bvh, _vertices(mesh), mesh.i0, mesh.i1, mesh.i2)
    got = orc.refit_moment_d64(bvh, rec, child, order)
    want = orc.refit_moment(bvh, order, dtype=np.float64)
    assert np.array_equal(got, want)


def test_moment_d64_root_normal_total_is_the_area_vector(orc):
    """Fields 4-6 (area-weighted normal N) are additive: the root's slots sum to 1/2 sum(e1 x e2)."""
    mesh = synth.open_soup(24, 12)
    bvh = orc.build_bvh4(mesh, "lbvh")
    vd = _perturbed(mesh)
    rec, child, _ = orc.double_tree(bvh, vd, mesh.i0, mesh.i1, mesh.i2)
    mom = orc.refit_moment_d64(bvh, rec, child, 0)
    r = rec.reshape(-1, 3, 3)
    area_vec = 0.5 * np.cross(r[:, 1] - r[:, 0], r[:, 2] - r[:, 0]).sum(axis=0)
    total = mom[0, 4:7, :].sum(axis=1)
    assert np.allclose(total, area_vec, rtol=0, atol=1e-13 * np.abs(area_vec).max())


def test_taylor_d64_descending_everywhere_is_the_exact_sum(orc):
    mesh = synth.displaced_sphere(32, 16)
    bvh = orc.build_bvh4(mesh, "hploc", 8)
    vd = _perturbed(mesh)
    rec, child, _ = orc.double_tree(bvh, vd, mesh.i0, mesh.i1, mesh.i2)
    q = [a.astype(np.float64) for a in synth.uniform_box_points(9, 200, (-1.5,) * 3, (1.5,) * 3)]
    exact = orc.winding_exact_mesh_d64(vd, mesh.i0, mesh.i1, mesh.i2, *q)
    for order in (0, 1, 2):
        mom = orc.refit_moment_d64(bvh, rec, child, order)
        w = orc.winding_taylor_d64(bvh, rec, mom, order, *q, accuracy_scale=1e6, stack_capacity=128)
        assert np.abs(w - exact).max() <= 1e-12


def test_taylor_d64_matches_float_replay_and_converges_with_order(orc):
    mesh = synth.displaced_sphere(32, 16)
    bvh = orc.build_bvh4(mesh, "hploc", 8)
    rec, child, _ = orc.double_tree(bvh, _vertices(mesh), mesh.i0, mesh.i1, mesh.i2)
    q = synth.uniform_box_points(10, 2000, (-8.0,) * 3, (8.0,) * 3)
    far = np.sqrt(q[0].astype(np.float64) ** 2 + q[1] ** 2 + q[2] ** 2) >= 4.0  # far field: the series converges
    q = [np.ascontiguousarray(a[far]) for a in q]
    qd = [a.astype(np.float64) for a in q]
    exact = orc.winding_exact_mesh_d64(_vertices(mesh), mesh.i0, mesh.i1, mesh.i2, *qd)
    err = []
    for order in (0, 1, 2):
        mom_d = orc.refit_moment_d64(bvh, rec, child, order)
        w = orc.winding_taylor_d64(bvh, rec, mom_d, order, *qd, accuracy_scale=2.0, stack_capacity=128)
        w32 = orc.winding_taylor_f32(bvh, orc.refit_moment(bvh, order), order, *q, accuracy_scale=2.0,
                                     stack_capacity=128)
        # Same decisions and terms; only the rounding of float differs.
        assert np.abs(w - w32).max() <= 2e-5
        err.append(np.sqrt(np.mean((w - exact) ** 2)))
    assert err[2] < err[1] < err[0]
