"""Differentiable winding numbers and unsigned distances for PyTorch (device tensors on ROCm).

A MeshScene holds the geometry, BVH and Taylor-moment objects of one triangle mesh whose topology
stays fixed while its vertex positions change.  Each forward call with a vertex tensor it has not
seen yet uploads the positions on the device, refits the BVH bounds and the moments (topology and
primitive order are kept), then runs the query batch.  Backward calls the adjoint batch of the
same library, which re-traverses the BVH with the forward's decisions instead of storing
per-query stacks, and returns dL/dV and dL/dq.

    scene = MeshScene(faces, vertices, order=1)
    w = taylor_winding(vertices, queries, scene)          # (Q,)
    d = unsigned_distance(vertices, points, scene)        # (P,)
    loss = ((torch.sigmoid(20 * (w - 0.5)) - labels) ** 2).mean() + d.mean()
    loss.backward()                                       # vertices.grad: (V, 3)

No counterpart in the reference, whose only derivative is the Taylor gradient with respect to the
query point (include/gwn/gwn_query.cuh:656-756).
"""

from __future__ import annotations

import numpy as np
import torch

from . import gwn


def _soa(t: torch.Tensor) -> torch.Tensor:
    """(N, 3) -> contiguous (3, N) float32 on the tensor's device."""
    return t.detach().to(torch.float32).t().contiguous()


class MeshScene:
    """Device state of one mesh with fixed topology (faces) and changing vertex positions."""

    def __init__(self, faces, vertices: torch.Tensor, order: int = 1,
                 method: gwn.BuildMethod = gwn.BuildMethod.HPLOC, hploc_search_radius: int = 8,
                 accuracy_scale: float = 2.0, culling_band: float = float("inf")):
        if order not in (0, 1):
            raise ValueError("the winding adjoint supports Taylor orders 0 and 1")
        f = np.asarray(faces.cpu() if isinstance(faces, torch.Tensor) else faces, dtype=np.int64).reshape(-1, 3)
        v = vertices.detach().cpu().to(torch.float32).numpy().reshape(-1, 3)
        self.geometry = gwn.GeometryObject()
        gwn.gwn_upload_geometry(self.geometry, v[:, 0], v[:, 1], v[:, 2],
                                f[:, 0].astype(np.uint32), f[:, 1].astype(np.uint32),
                                f[:, 2].astype(np.uint32)).raise_if_error()
        self.bvh = gwn.BvhObject()
        gwn.gwn_build_bvh(self.geometry, self.bvh, gwn.BuildOptions(
            method=method, hploc_search_radius=hploc_search_radius)).raise_if_error()
        self.moment = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(self.bvh, self.moment).raise_if_error()
        self.vertex_count = v.shape[0]
        self.accuracy_scale = float(accuracy_scale)
        self.culling_band = float(culling_band)
        self._loaded = (vertices.data_ptr(), vertices._version)
        self.stamp = 0

    def update(self, vertices: torch.Tensor) -> None:
        """Load new positions (refit bounds and moments) unless this tensor state is loaded."""
        key = (vertices.data_ptr(), vertices._version)
        if key == self._loaded:
            return
        if vertices.shape != (self.vertex_count, 3):
            raise ValueError(f"vertices must have shape ({self.vertex_count}, 3)")
        xyz = _soa(vertices)
        gwn.gwn_update_geometry(self.geometry, xyz[0], xyz[1], xyz[2]).raise_if_error()
        gwn.gwn_refit_bvh(self.geometry, self.bvh).raise_if_error()
        gwn.gwn_refit_bvh_moment(self.bvh, self.moment).raise_if_error()
        self._loaded = key
        self.stamp += 1


class TaylorWinding(torch.autograd.Function):
    """w_q = Taylor winding number of the scene's mesh at query q; differentiable in both."""

    @staticmethod
    def forward(ctx, vertices, queries, scene: MeshScene):
        scene.update(vertices)
        q = _soa(queries)
        w = torch.empty(q.shape[1], dtype=torch.float32, device=queries.device)
        gwn.gwn_compute_winding_number_taylor_batch(scene.bvh, scene.moment, q[0], q[1], q[2], w,
                                                    accuracy_scale=scene.accuracy_scale).raise_if_error()
        ctx.scene, ctx.stamp = scene, scene.stamp
        ctx.save_for_backward(q)
        return w

    @staticmethod
    def backward(ctx, grad_w):
        (q,) = ctx.saved_tensors
        scene = ctx.scene
        if scene.stamp != ctx.stamp:
            raise RuntimeError("mesh positions changed between forward and backward")
        gv = torch.zeros(3, scene.vertex_count, dtype=torch.float32, device=q.device)
        want_q = ctx.needs_input_grad[1]
        gq = torch.empty(3, q.shape[1], dtype=torch.float32, device=q.device) if want_q else None
        gwn.gwn_compute_winding_taylor_adjoint_batch(
            scene.geometry, scene.bvh, scene.moment, q[0], q[1], q[2], grad_w.to(torch.float32).contiguous(),
            gv[0], gv[1], gv[2], *(gq if want_q else (None, None, None)),
            accuracy_scale=scene.accuracy_scale).raise_if_error()
        return gv.t(), (gq.t() if want_q else None), None


class UnsignedDistance(torch.autograd.Function):
    """d_p = distance from point p to the scene's mesh; differentiable in both."""

    @staticmethod
    def forward(ctx, vertices, points, scene: MeshScene):
        scene.update(vertices)
        p = _soa(points)
        d = torch.empty(p.shape[1], dtype=torch.float32, device=points.device)
        # The nearest triangle of every point is kept for backward, which then needs no traversal.
        nearest = torch.empty(p.shape[1], dtype=torch.int32, device=points.device)
        gwn.gwn_compute_unsigned_distance_batch(scene.bvh, p[0], p[1], p[2], d, culling_band=scene.culling_band,
                                                nearest=nearest).raise_if_error()
        ctx.scene, ctx.stamp = scene, scene.stamp
        ctx.save_for_backward(p, nearest)
        return d

    @staticmethod
    def backward(ctx, grad_d):
        p, nearest = ctx.saved_tensors
        scene = ctx.scene
        if scene.stamp != ctx.stamp:
            raise RuntimeError("mesh positions changed between forward and backward")
        gv = torch.zeros(3, scene.vertex_count, dtype=torch.float32, device=p.device)
        want_p = ctx.needs_input_grad[1]
        gp = torch.empty(3, p.shape[1], dtype=torch.float32, device=p.device) if want_p else None
        gwn.gwn_compute_unsigned_distance_adjoint_batch(
            scene.geometry, scene.bvh, p[0], p[1], p[2], grad_d.to(torch.float32).contiguous(), gv[0], gv[1], gv[2],
            *(gp if want_p else (None, None, None)), nearest=nearest).raise_if_error()
        return gv.t(), (gp.t() if want_p else None), None


def taylor_winding(vertices: torch.Tensor, queries: torch.Tensor, scene: MeshScene) -> torch.Tensor:
    return TaylorWinding.apply(vertices, queries, scene)


def unsigned_distance(vertices: torch.Tensor, points: torch.Tensor, scene: MeshScene) -> torch.Tensor:
    return UnsignedDistance.apply(vertices, points, scene)
