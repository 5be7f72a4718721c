"""torch.autograd wrappers (smallgwn_amd.autograd) on the differentiation workload's loss, checked
against the FP64 adjoint oracle: L = mean((sigmoid(k (w - 0.5)) - label)^2) + mean(d)."""

import numpy as np
import pytest

from smallgwn_amd import synth

pytestmark = pytest.mark.gpu


def _loss_parts(torch, w, d, labels, k):
    s = torch.sigmoid(k * (w - 0.5))
    return ((s - labels) ** 2).mean() + d.mean()


@pytest.mark.parametrize("order", [0, 1])
def test_autograd_loss_gradient_matches_oracle(gwn, orc, dev, order):
    import torch

    from smallgwn_amd import autograd as ag

    mesh = synth.displaced_sphere(48, 24)
    faces = np.stack([mesh.i0, mesh.i1, mesh.i2], 1)
    v0 = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    vertices = torch.tensor(v0, device=dev, requires_grad=True)
    scene = ag.MeshScene(faces, vertices, order=order, method=gwn.BuildMethod.LBVH)
    qx, qy, qz = synth.uniform_box_points(3, 8192, (-1.3,) * 3, (1.3,) * 3)
    queries = torch.tensor(np.stack([qx, qy, qz], 1), device=dev, requires_grad=True)
    labels = (queries.detach().norm(dim=1) < 0.9).to(torch.float32)
    px, py, pz = synth.uniform_box_points(4, 4096, (-1.2,) * 3, (1.2,) * 3)
    points = torch.tensor(np.stack([px, py, pz], 1), device=dev)
    k = 20.0
    w = ag.taylor_winding(vertices, queries, scene)
    d = ag.unsigned_distance(vertices, points, scene)
    loss = _loss_parts(torch, w, d, labels, k)
    loss.backward()
    gv = vertices.grad.cpu().numpy().T.astype(np.float64)
    gq = queries.grad.cpu().numpy().astype(np.float64)

    # Oracle: upstream gradients from the same forward values, then the FP64 adjoints.
    wn = w.detach().cpu().numpy().astype(np.float64)
    s = 1.0 / (1.0 + np.exp(-k * (wn - 0.5)))
    lab = labels.cpu().numpy().astype(np.float64)
    uw = 2.0 * (s - lab) * s * (1.0 - s) * k / wn.shape[0]
    ud = np.full(px.shape[0], 1.0 / px.shape[0])
    rb = orc.build_bvh4(mesh)
    rdv_w, rdq = orc.winding_taylor_adjoint_f64(rb, orc.refit_moment(rb, order), order, 2.0, mesh, qx, qy, qz, uw)
    _, rdv_d, _ = orc.distance_adjoint_f64(mesh, px, py, pz, ud)
    ref = rdv_w + rdv_d
    assert np.abs(gv - ref).max() <= 1e-3 * np.abs(ref).max(), np.abs(gv - ref).max() / np.abs(ref).max()
    assert np.abs(gq - rdq).max() <= 1e-3 * (1e-12 + np.abs(rdq).max())


def test_scene_refits_once_per_vertex_state_and_detects_stale_backward(gwn, orc, dev):
    import torch

    from smallgwn_amd import autograd as ag

    mesh = synth.icosphere(2)
    faces = np.stack([mesh.i0, mesh.i1, mesh.i2], 1)
    v = torch.tensor(np.stack([mesh.vx, mesh.vy, mesh.vz], 1), device=dev, requires_grad=True)
    scene = ag.MeshScene(faces, v, order=1)  # H-PLOC tree, search radius 8
    q = torch.tensor([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]], device=dev)
    w = ag.taylor_winding(v, q, scene)
    rb = orc.build_bvh4(mesh, "hploc", 8)
    ref = orc.winding_taylor_f32(rb, orc.refit_moment(rb, 1), 1, np.float32([0, 2]), np.float32([0, 0]),
                                 np.float32([0, 0]), accuracy_scale=2.0)
    got = w.detach().cpu().numpy()
    assert scene.stamp == 0 and np.all(np.abs(got - ref) <= 1e-6 * (1 + np.abs(ref)))
    # A new vertex state (the sphere scaled by 1/2, exact in FP32) is refit once; the Taylor
    # winding at the centre is scale invariant, and the stale graph refuses backward.
    v2 = (v * 0.5).detach().requires_grad_(True)
    w2 = ag.taylor_winding(v2, q, scene)
    assert scene.stamp == 1 and abs(float(w2[0]) - float(w[0])) <= 1e-6
    with pytest.raises(RuntimeError):
        w.sum().backward()
    w2.sum().backward()
    assert torch.isfinite(v2.grad).all()
