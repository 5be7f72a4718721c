"""GPU parity of the ray first-hit batch (gwn_compute_ray_first_hit_batch).

The device runs the traversal the reference's launcher selects (packed near-first when
packed_stack_bound fits the stack, canonical otherwise) with unfused FP32 arithmetic, so every
32-byte hit record -- t, barycentrics, normal, primitive id, status -- equals the oracle's
replay bit for bit, ties included.
"""

import numpy as np
import pytest

from smallgwn_amd import synth
from tests.test_gpu_build import MESHES, build_gpu

pytestmark = pytest.mark.gpu


def _t(a, dev):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)


def _rays(gwn, dev, bvh, o, d, t_min=0.0, t_max=float("inf"), stack=64):
    import torch

    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    n = o.shape[0]
    out = torch.full((8 * n,), -123, dtype=torch.int32, device=dev)
    gwn.gwn_compute_ray_first_hit_batch(bvh, *(_t(o[:, k], dev) for k in range(3)), *(_t(d[:, k], dev) for k in range(3)),
                                        out, t_min, t_max, config=gwn.QueryConfig(stack_capacity=stack)).raise_if_error()
    torch.cuda.synchronize()
    return out.cpu().numpy().view(gwn.RAY_HIT_DTYPE)


def _oracle(orc, mesh, o, d, t_min=0.0, t_max=float("inf"), stack=64):
    o = np.asarray(o, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(d, dtype=np.float32).reshape(-1, 3)
    return orc.ray_first_hit_f32(orc.build_bvh4(mesh), *(o[:, k] for k in range(3)), *(d[:, k] for k in range(3)),
                                 t_min, t_max, stack_capacity=stack)


def _same(got, ref):
    assert got.tobytes() == ref.tobytes(), (got[:4], ref[:4])


def test_ray_reference_kats(gwn, orc, dev):
    """tests/unit_gwn_query_distance_ray.cu:222-298, 350-410, 465-482."""
    _, one = build_gpu(gwn, synth.layered_triangles(1))
    h = _rays(gwn, dev, one, [0.25, 0.25, 1], [0, 0, -1])[0]
    assert h["status"] == 1 and h["primitive_id"] == 0 and h["t"] == 1.0 and h["nz"] == 1.0
    miss = _rays(gwn, dev, one, [0.25, 0.25, 1], [0, 0, 0], 0.0, 10.0)[0]
    assert miss["status"] == 0 and miss["t"] == -1.0 and miss["primitive_id"] == 0xFFFFFFFF
    _, twelve = build_gpu(gwn, synth.layered_triangles(12))
    h = _rays(gwn, dev, twelve, [[0.25, 0.25, 20], [2, 2, 20], [0, 0.25, 20]], [[0, 0, -1]] * 3)
    assert h["t"].tolist() == [9.0, -1.0, 9.0] and h["primitive_id"].tolist() == [11, 0xFFFFFFFF, 11]
    _, sixteen = build_gpu(gwn, synth.layered_triangles(16))
    h = _rays(gwn, dev, sixteen, [0.25, 0.25, 20], [0, 0, 1], -30.0, 0.0)[0]
    assert h["status"] == 1 and h["primitive_id"] == 0 and h["t"] == -20.0
    octa = synth._mesh_from_arrays(
        np.stack([[1, -1, 0, 0, 0, 0], [0, 0, 1, -1, 0, 0], [0, 0, 0, 0, 1, -1]], 1),
        np.stack([[0, 2, 1, 3, 2, 1, 3, 0], [2, 1, 3, 0, 0, 2, 1, 3], [4, 4, 4, 4, 5, 5, 5, 5]], 1))
    _, ob = build_gpu(gwn, octa)
    o, d = [[2, 2, 2], [2, 0.2, 0.1], [2, 2, 2]], [[-1, -1, -1], [-1, 0, 0], [1, 1, 1]]
    h = _rays(gwn, dev, ob, o, d)
    assert abs(h["t"][0] - 5 / 3) <= 1e-5 and abs(h["t"][1] - 1.3) <= 1e-5 and h["t"][2] == -1.0
    _same(h, _oracle(orc, octa, o, d))


@pytest.mark.parametrize("name", ["icosphere4", "torus", "soup", "displaced_256x128", "coincident33"])
def test_ray_matches_oracle_bitwise(gwn, orc, dev, name):
    mesh = MESHES[name]
    _, bvh = build_gpu(gwn, mesh)
    rng = np.random.default_rng(3)
    n = 4096
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 3.0
    d = rng.uniform(-1, 1, size=(n, 3)) - o
    d[::97, 0] = 0.0  # some rays with a parallel axis
    for t_min, t_max in ((0.0, float("inf")), (0.5, 2.5)):
        _same(_rays(gwn, dev, bvh, o, d, t_min, t_max), _oracle(orc, mesh, o, d, t_min, t_max))


def test_ray_canonical_traversal_when_packed_bound_exceeds_stack(gwn, orc, dev):
    """A stack capacity between internal_stack_bound and packed_stack_bound selects the canonical
    traversal (gwn_query.cuh:281-289); the records still match the oracle's same selection."""
    mesh = MESHES["displaced_256x128"]
    _, bvh = build_gpu(gwn, mesh)
    info = bvh.info()
    cap = info["internal_stack_bound"]
    assert cap < info["packed_stack_bound"]
    rng = np.random.default_rng(9)
    o = rng.normal(size=(2048, 3)) * 2
    d = -o + rng.normal(size=(2048, 3)) * 0.1
    _same(_rays(gwn, dev, bvh, o, d, stack=cap), _oracle(orc, mesh, o, d, stack=cap))


def test_ray_large_mesh_subsample(gwn, orc, dev):
    mesh = synth.config_large_mesh()
    _, bvh = build_gpu(gwn, mesh)
    r = synth.rays_toward_ball(65536, seed=7)
    o = np.stack(r[0:3], 1)
    d = np.stack(r[3:6], 1)
    got = _rays(gwn, dev, bvh, o, d)
    idx = np.arange(0, 65536, 16)
    _same(got[idx], _oracle(orc, mesh, o[idx], d[idx]))
    assert (got["status"] == 1).mean() > 0.5


def test_ray_validation_contract(gwn, dev):
    """gwn_query.cuh:245-262."""
    import torch

    _, bvh = build_gpu(gwn, MESHES["torus"])
    r = [torch.zeros(4, device=dev) for _ in range(6)]
    out = torch.zeros(32, dtype=torch.int32, device=dev)
    st = gwn.gwn_compute_ray_first_hit_batch(gwn.BvhObject(), *r, out)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "no queryable data" in st.message
    st = gwn.gwn_compute_ray_first_hit_batch(bvh, *r, out, t_min=1.0, t_max=0.5)
    assert st.error == gwn.Error.INVALID_ARGUMENT and "t_max >= t_min" in st.message
    st = gwn.gwn_compute_ray_first_hit_batch(bvh, *r, torch.zeros(8, dtype=torch.int32, device=dev))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "mismatched span sizes" in st.message
    st = gwn.gwn_compute_ray_first_hit_batch(bvh, *r, out, config=gwn.QueryConfig(stack_capacity=1))
    assert st.error == gwn.Error.INVALID_ARGUMENT and "Traversal stack capacity" in st.message
