"""GPU parity of the boundary chain (gwn_build_boundary_chain): rows equal the oracle's exactly,
in the same (min vertex, max vertex) order."""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES

pytestmark = pytest.mark.gpu


def _build(gwn, mesh):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    b = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, b).raise_if_error()
    return b


def test_boundary_kats(gwn):
    """tests/unit_gwn_boundary.cu:94-187."""
    empty = _build(gwn, synth._mesh_from_arrays(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64)))
    assert empty.has_data() and empty.info()["edge_count"] == 0
    closed = _build(gwn, synth.closed_cube())
    info = closed.info()
    assert info["is_built"] and info["edge_count"] == 0 and info["mesh_triangle_count"] == 12
    s, e, m = _build(gwn, synth.open_cube()).download()
    assert sorted(zip(s.tolist(), e.tolist(), m.tolist())) == [(4, 7, 1), (5, 4, 1), (6, 5, 1), (7, 6, 1)]
    net = synth._mesh_from_arrays(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]),
                                  np.stack([[0, 1, 0, 0], [1, 0, 1, 1], [2, 3, 2, 2]], 1))
    rows = list(zip(*(a.tolist() for a in _build(gwn, net).download())))
    assert (0, 1, 2) in rows and not any(r[0] == 1 and r[1] == 0 for r in rows)


@pytest.mark.parametrize("name", ["open_soup", "soup", "torus", "coincident33", "large_open"])
def test_boundary_matches_oracle(gwn, orc, name):
    if name == "open_soup":
        mesh = synth.open_soup(128, 64, seed=3)
    elif name == "large_open":
        mesh = synth.open_soup(512, 256, seed=4)
    else:
        mesh = MESHES[name]
    got = _build(gwn, mesh).download()
    ref = orc.boundary_chain(mesh)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_boundary_rebuild_replaces_and_validates(gwn):
    b = _build(gwn, synth.open_cube())
    assert b.info()["edge_count"] == 4
    geom = gwn.GeometryObject()
    m = synth.closed_cube()
    gwn.gwn_upload_geometry(geom, m.vx, m.vy, m.vz, m.i0, m.i1, m.i2).raise_if_error()
    gwn.gwn_build_boundary_chain(geom, b).raise_if_error()
    assert b.info()["edge_count"] == 0
