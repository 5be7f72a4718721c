"""PLY reader (smallgwn_amd/ply.py) on the reference's fixture meshes and on written files.

The fixtures in tests/golden/meshes are the reference's tests/data/*.ply, kept as data.  Their
contents are pinned by what the reference's own tests state about them (8 vertices and 12 or 10
triangles, closed or open) and by the boundary-chain known answer for the open cube
(tests/unit_gwn_boundary.cu:94-187: rows (4,7,1), (5,4,1), (6,5,1), (7,6,1)), computed by the
oracle from the loaded topology.  Written files round-trip bit for bit in ASCII and both binary
byte orders.
"""

import os

import numpy as np
import pytest

from smallgwn_amd import ply, synth

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def test_collects_the_fixture_meshes_in_sorted_order():
    names = [os.path.basename(p) for p in ply.collect_ply_model_paths(MESHES)]
    assert names == ["closed_cube.ply", "nonplanar_closed.ply", "open_cube.ply"]


@pytest.mark.parametrize("name,tris", [("closed_cube", 12), ("nonplanar_closed", 12), ("open_cube", 10)])
def test_fixture_counts_and_coordinates(name, tris):
    mesh = ply.load_ply_mesh(os.path.join(MESHES, name + ".ply"))
    assert mesh is not None
    assert (mesh.vertex_count, mesh.triangle_count) == (8, tris)
    assert mesh.vx.dtype == np.float32 and mesh.i0.dtype == np.uint32
    # First vertex (-1, -1, -1) and first face (0, 2, 1) in every fixture.
    assert (mesh.vx[0], mesh.vy[0], mesh.vz[0]) == (-1.0, -1.0, -1.0)
    assert (mesh.i0[0], mesh.i1[0], mesh.i2[0]) == (0, 2, 1)


def test_closed_fixtures_are_closed_and_open_cube_has_the_reference_boundary(orc):
    for name in ("closed_cube", "nonplanar_closed"):
        s, e, m = orc.boundary_chain(ply.load_ply_mesh(os.path.join(MESHES, name + ".ply")))
        assert s.size == 0
    s, e, m = orc.boundary_chain(ply.load_ply_mesh(os.path.join(MESHES, "open_cube.ply")))
    assert sorted(zip(s.tolist(), e.tolist(), m.tolist())) == [(4, 7, 1), (5, 4, 1), (6, 5, 1), (7, 6, 1)]


def test_closed_cube_winding_is_one_inside_zero_outside(orc):
    mesh = ply.load_ply_mesh(os.path.join(MESHES, "closed_cube.ply"))
    w = orc.winding_exact_mesh_f64(mesh, np.array([0.1, 3.0], np.float32), np.array([-0.2, 0.0], np.float32),
                                   np.array([0.3, 0.0], np.float32))
    assert abs(w[0] - 1.0) < 1e-12 and abs(w[1]) < 1e-12


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip(tmp_path, binary):
    mesh = synth.torus(7, 5)
    path = str(tmp_path / "m.ply")
    ply.write_ply(path, mesh, binary=binary)
    back = ply.load_ply_mesh(path)
    for a, b in zip((mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2),
                    (back.vx, back.vy, back.vz, back.i0, back.i1, back.i2)):
        assert np.array_equal(a, b)


def _write(path, text, body=b""):
    with open(path, "wb") as f:
        f.write(text.encode("ascii"))
        f.write(body)


def test_big_endian_double_with_extra_element_and_properties(tmp_path):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    header = ("ply\nformat binary_big_endian 1.0\ncomment test\nelement vertex 4\nproperty double x\n"
              "property double y\nproperty double z\nproperty uchar red\nelement edge 1\nproperty int a\n"
              "property int b\nelement face 2\nproperty uchar flags\nproperty list uchar int vertex_indices\n"
              "end_header\n")
    vrec = np.zeros(4, dtype=[("x", ">f8"), ("y", ">f8"), ("z", ">f8"), ("r", "u1")])
    vrec["x"], vrec["y"], vrec["z"], vrec["r"] = v[:, 0], v[:, 1], v[:, 2], 7
    erec = np.array([(0, 1)], dtype=[("a", ">i4"), ("b", ">i4")])
    frec = np.array([(1, 3, (0, 1, 2)), (0, 3, (0, 3, 1))], dtype=[("f", "u1"), ("n", "u1"), ("i", ">i4", (3,))])
    path = str(tmp_path / "be.ply")
    _write(path, header, vrec.tobytes() + erec.tobytes() + frec.tobytes())
    mesh = ply.load_ply_mesh(path)
    assert mesh.vertex_count == 4 and mesh.triangle_count == 2
    assert np.array_equal(np.stack([mesh.i0, mesh.i1, mesh.i2], 1), [[0, 1, 2], [0, 3, 1]])
    assert np.array_equal(mesh.vz, [0, 0, 0, 1])


def test_polygons_split_into_fans_in_face_order(tmp_path):
    text = ("ply\nformat ascii 1.0\nelement vertex 6\nproperty float x\nproperty float y\nproperty float z\n"
            "element face 2\nproperty list uchar uint vertex_indices\nend_header\n"
            "0 0 0\n1 0 0\n1 1 0\n0 1 0\n2 0 0\n2 1 0\n4 0 1 2 3\n3 1 4 5\n")
    path = str(tmp_path / "quad.ply")
    _write(path, text)
    mesh = ply.load_ply_mesh(path)
    assert np.array_equal(np.stack([mesh.i0, mesh.i1, mesh.i2], 1), [[0, 1, 2], [0, 2, 3], [1, 4, 5]])
    # The binary form of mixed face sizes takes the record-at-a-time path.
    body = np.array([0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, 2, 0, 0, 2, 1, 0], "<f4").tobytes()
    body += bytes([4]) + np.array([0, 1, 2, 3], "<u4").tobytes() + bytes([3]) + np.array([1, 4, 5], "<u4").tobytes()
    bpath = str(tmp_path / "quadb.ply")
    _write(bpath, text.split("end_header\n")[0].replace("ascii", "binary_little_endian") + "end_header\n", body)
    bm = ply.load_ply_mesh(bpath)
    assert np.array_equal(np.stack([bm.i0, bm.i1, bm.i2], 1), [[0, 1, 2], [0, 2, 3], [1, 4, 5]])


@pytest.mark.parametrize("body,why", [
    ("0 0 0\n1 0 0\nnan 1 0\n3 0 1 2\n", "non-finite"),
    ("0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n", "index out of range"),
    ("0 0 0\n1 0 0\n0 1 0\n2 0 1\n", "no triangle"),
])
def test_unusable_meshes_are_skipped(tmp_path, body, why):
    text = ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
            "element face 1\nproperty list uchar int vertex_indices\nend_header\n" + body)
    path = str(tmp_path / "bad.ply")
    _write(path, text)
    assert ply.load_ply_mesh(path) is None, why


def test_not_a_ply_path_and_malformed_files(tmp_path):
    assert ply.load_ply_mesh(str(tmp_path / "mesh.obj")) is None
    path = str(tmp_path / "trunc.ply")
    _write(path, "ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                 "property float z\nend_header\n", b"\0" * 20)
    with pytest.raises(ply.PlyError):
        ply.read_ply(path)
    _write(path, "plx\n")
    with pytest.raises(ply.PlyError):
        ply.read_ply(path)
