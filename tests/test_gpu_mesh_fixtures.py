"""The reference's mesh-directory integration contracts (builder, static, dynamic) on its fixture
meshes (tests/golden/meshes, the reference's tests/data/*.ply) and on synthetic meshes written to
binary PLY and read back, through the C ABI on the GPU.  The contracts themselves are in
tests/mesh_e2e.py."""

import os

import pytest

from smallgwn_amd import ply, synth
from tests import mesh_e2e

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")
FIXTURES = ["closed_cube", "nonplanar_closed", "open_cube"]


def _load(name):
    mesh = ply.load_ply_mesh(os.path.join(MESHES, name + ".ply"))
    assert mesh is not None
    return mesh


@pytest.mark.parametrize("name", FIXTURES)
def test_builder_contract_on_fixture(gwn, orc, dev, name):
    print(f"[mesh] builder {name}", flush=True)
    mesh_e2e.builder_contract(gwn, orc, dev, _load(name), radii=(1, 4, 8))


@pytest.mark.parametrize("name", FIXTURES)
def test_static_ray_contract_on_fixture(gwn, dev, name):
    print(f"[mesh] static {name}", flush=True)
    mesh_e2e.static_contract(gwn, dev, _load(name))


@pytest.mark.parametrize("name", FIXTURES)
def test_dynamic_refit_contract_on_fixture(gwn, dev, name):
    print(f"[mesh] dynamic {name}", flush=True)
    mesh_e2e.dynamic_contract(gwn, dev, _load(name))


def test_directory_runner_on_written_meshes(gwn, orc, dev, tmp_path):
    """The dataset runner over a directory: the fixtures plus written torus and displaced-sphere
    meshes (binary PLY), and a file it must skip."""
    print("[mesh] directory runner", flush=True)
    for name in FIXTURES:
        with open(os.path.join(MESHES, name + ".ply"), "rb") as src, open(tmp_path / (name + ".ply"), "wb") as dst:
            dst.write(src.read())
    ply.write_ply(str(tmp_path / "torus.ply"), synth.torus(24, 12), binary=True)
    ply.write_ply(str(tmp_path / "sphere.ply"), synth.displaced_sphere(32, 16), binary=False)
    (tmp_path / "empty.ply").write_bytes(b"ply\nformat ascii 1.0\nelement vertex 0\nproperty float x\n"
                                         b"property float y\nproperty float z\nelement face 0\n"
                                         b"property list uchar int vertex_indices\nend_header\n")
    skipped = []
    assert mesh_e2e.run_directory(gwn, orc, dev, str(tmp_path), log=skipped.append) == 5
    assert any("failed to parse" in s and "empty.ply" in s for s in skipped)
