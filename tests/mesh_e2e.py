"""Dataset end-to-end contracts on PLY meshes, through the C ABI on the GPU.

Restates the reference's mesh-directory integration suites (run there with ``--mesh-dir``,
tests/test_mesh_main.cpp:10-40) for each mesh a directory holds:

- builder (tests/integration_gwn_builder.cu:36-119, 120-298): LBVH and H-PLOC (radius 8) trees
  (width 4 unless asked otherwise) whose internal nodes are complete (every node reached once, leaf ranges in bounds, stack
  bounds equal the host recursion), then the exact batch and a Taylor order-2 batch at accuracy
  scale 1e6 (every finite child descends) on nine box-derived points against the FP64 exact
  winding number, within 2e-5 + 2e-5 max(|a|, |b|);
- static (tests/integration_gwn_static.cu:44-356): two centroid rays per sampled triangle and 32
  Fibonacci-sphere rays against an independent FP64 Moller-Trumbore scan: equal hit status and t
  within 5e-5 max(1, |t|), on both builders;
- dynamic (tests/integration_gwn_dynamic.cu:66-426): deform the mesh by a fixed affine map,
  update positions, refit bounds and order-2 moments, and compare distance, Taylor winding and
  gradient, antipodal winding and gradient and first hits with a fresh build of the deformed mesh
  (the reference's tolerances; hit normals must face the triangle's own normal).

Run as ``python -m tests.mesh_e2e --mesh-dir DIR`` (a GPU and the built library are needed); the
GPU test suite runs the same contracts on the reference's fixture meshes in tests/golden/meshes.
"""

from __future__ import annotations

import argparse
import sys

import numpy as np

STACK = 64  # the reference tests' traversal stack capacity (tests/test_utils.cuh)


def _dev_arrays(dev, *arrays):
    import torch

    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrays]


def _upload(gwn, mesh):
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    return geom


def _build(gwn, geom, method, radius=8, width=4):
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=method, hploc_search_radius=radius,
                                                  width=width)).raise_if_error()
    return bvh


# ---- builder contract ----------------------------------------------------------------------------

def check_internal_nodes_complete(bvh, triangle_count: int) -> None:
    """tests/integration_gwn_builder.cu:36-119 on the downloaded node array."""
    info = bvh.info()
    raw, prim, _ = bvh.download()
    width = info["width"]
    child = np.dtype([("bounds", np.float32, (6,)), ("reference", np.uint64)])
    nodes = np.frombuffer(raw.tobytes(), dtype=np.dtype([("children", child, (width,))]))
    root = int(info["root_reference"])
    valid_bit, count_shift = 1 << 47, 48
    if (root >> count_shift) != 0:  # leaf root
        assert info["node_count"] == 0
        assert info["internal_stack_bound"] == 0 and info["packed_stack_bound"] == 0
        return
    assert root & valid_bit, "root is neither internal nor leaf"
    n = nodes.shape[0]
    visited = np.zeros(n, dtype=bool)
    reach = 0

    def bounds(i):
        nonlocal reach
        assert i < n, "internal child offset out of bounds"
        assert not visited[i], "internal child reachable more than once"
        visited[i] = True
        reach += 1
        valid = internal = 0
        mi = mp = 0
        for c in nodes[i]["children"]:
            ref = int(c["reference"])
            if not ref & valid_bit:
                continue
            valid += 1
            off, cnt = ref & ((1 << 47) - 1), ref >> count_shift
            if cnt == 0:
                internal += 1
                ci, cp = bounds(off)
                mi, mp = max(mi, ci), max(mp, cp)
            else:
                assert off <= prim.shape[0] and cnt <= prim.shape[0] - off, "leaf range out of bounds"
        assert valid > 0, "internal node has no valid children"
        return (internal - 1 + mi if internal else 0), valid - 1 + mp

    ib, pb = bounds(root & ((1 << 47) - 1))
    assert reach == n
    assert info["internal_stack_bound"] == ib and info["packed_stack_bound"] == pb
    assert sorted(prim.tolist()) == list(range(triangle_count))


def builder_queries(mesh):
    """The nine points of tests/integration_gwn_builder.cu:120-168 (float arithmetic)."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    lo, hi = v.min(0), v.max(0)
    c = (lo + hi) * np.float32(0.5)
    e = np.maximum(hi - lo, np.float32(1e-2))
    f = np.float32
    pts = [c.copy() for _ in range(9)]
    for k, (axis, sgn) in enumerate(((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1))):
        pts[1 + k][axis] = c[axis] + f(sgn) * f(1.3) * e[axis]
    pts[7] = np.array([c[0] + f(0.31) * e[0], c[1] - f(0.27) * e[1], c[2] + f(0.23) * e[2]], np.float32)
    pts[8] = np.array([c[0] - f(0.19) * e[0], c[1] + f(0.37) * e[1], c[2] - f(0.29) * e[2]], np.float32)
    p = np.stack(pts).astype(np.float32)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()


def builder_contract(gwn, orc, dev, mesh, methods=None, radii=(8,), queries=None, width: int = 4) -> None:
    import torch

    q = queries if queries is not None else builder_queries(mesh)
    ref = orc.winding_exact_mesh_f64(mesh, *q)
    d = _dev_arrays(dev, *q)
    geom = _upload(gwn, mesh)
    for method in methods or (gwn.BuildMethod.LBVH, gwn.BuildMethod.HPLOC):
        for radius in radii:
            bvh = _build(gwn, geom, method, radius, width)
            assert bvh.info()["width"] == width
            check_internal_nodes_complete(bvh, mesh.triangle_count)
            mom = gwn.MomentObject(2)
            gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
            n = q[0].shape[0]
            exact = torch.empty(n, dtype=torch.float32, device=dev)
            desc = torch.empty(n, dtype=torch.float32, device=dev)
            cfg = gwn.QueryConfig(stack_capacity=STACK)
            gwn.gwn_compute_winding_number_exact_batch(bvh, *d, exact, config=cfg).raise_if_error()
            gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, desc, accuracy_scale=1e6,
                                                        config=cfg).raise_if_error()
            torch.cuda.synchronize()
            ex, de = exact.cpu().numpy().astype(np.float64), desc.cpu().numpy().astype(np.float64)

            def tol(a, b):
                return 2e-5 + 2e-5 * np.maximum(np.abs(a), np.abs(b))

            assert np.all(np.abs(ex - ref) <= tol(ex, ref)), (method, radius, ex, ref)
            assert np.all(np.abs(de - ref) <= tol(de, ref)), (method, radius, de, ref)
            assert np.all(np.abs(de - ex) <= tol(de, ex)), (method, radius, de, ex)


# ---- static contract -----------------------------------------------------------------------------

def fixture_rays(mesh):
    """tests/integration_gwn_static.cu:44-174: centroid rays from both sides of up to 8 sampled
    triangles plus 32 Fibonacci-sphere rays (one quarter pointing away), directions unit in FP64
    and stored as float."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float64)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.int64)
    lo, hi = v.min(0), v.max(0)
    center = 0.5 * (lo + hi)
    extent = np.maximum(hi - lo, 1e-3)
    diag = float(np.sqrt((extent * extent).sum()))
    org, dirs = [], []

    def append(o, dvec):
        length = float(np.sqrt((dvec * dvec).sum()))
        if not length > 0.0:
            return
        org.append(o.astype(np.float32))
        dirs.append((dvec / length).astype(np.float32))

    T = f.shape[0]
    sampled = min(8, T)
    for s in range(sampled):
        t = s * T // sampled
        a, b, c = v[f[t, 0]], v[f[t, 1]], v[f[t, 2]]
        nrm = np.cross(b - a, c - a)
        nl = float(np.sqrt((nrm * nrm).sum()))
        if not nl > 0.0:
            continue
        cen = (a + b + c) / 3.0
        for side in (-1.0, 1.0):
            dvec = -side * nrm / nl
            append(cen - dvec * (0.75 * diag), dvec)
    golden = 2.39996322972865332
    for r in range(32):
        z = 1.0 - 2.0 * (r + 0.5) / 32
        rr = np.sqrt(max(0.0, 1.0 - z * z))
        ang = golden * r
        radial = np.array([rr * np.cos(ang), rr * np.sin(ang), z])
        o = center + 2.5 * diag * radial
        target = np.array([center[k] + 0.31 * extent[k] * np.sin(ang * (k + 1) + 0.37 * k) for k in range(3)])
        append(o, radial if r % 4 == 3 else target - o)
    o = np.stack(org)
    dv = np.stack(dirs)
    return tuple(o[:, k].copy() for k in range(3)) + tuple(dv[:, k].copy() for k in range(3))


def moller_trumbore_f64(mesh, ox, oy, oz, dx, dy, dz):
    """Independent FP64 first-hit t per ray (-1 on a miss), scanning every triangle
    (tests/integration_gwn_static.cu:176-258)."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float64)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.int64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    e1, e2 = b - a, c - a
    out = np.full(ox.shape[0], -1.0)
    for r in range(ox.shape[0]):
        o = np.array([ox[r], oy[r], oz[r]], np.float64)
        dvec = np.array([dx[r], dy[r], dz[r]], np.float64)
        p = np.cross(np.broadcast_to(dvec, e2.shape), e2)
        det = (e1 * p).sum(1)
        ok = np.abs(det) > 1e-12
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = o - a
        u = (tv * p).sum(1) * inv
        ok &= (u >= -1e-10) & (u <= 1.0 + 1e-10)
        qv = np.cross(tv, e1)
        vv = (dvec * qv).sum(1) * inv
        ok &= (vv >= -1e-10) & (u + vv <= 1.0 + 1e-10)
        t = (e2 * qv).sum(1) * inv
        ok &= t >= 0.0
        if ok.any():
            out[r] = t[ok].min()
    return out


def static_contract(gwn, dev, mesh) -> None:
    import torch

    rays = fixture_rays(mesh)
    n = rays[0].shape[0]
    assert n > 24
    ref_t = moller_trumbore_f64(mesh, *rays)
    hit = ref_t >= 0.0
    assert hit.sum() > n // 4 and (n - hit.sum()) > n // 8
    d = _dev_arrays(dev, *rays)
    geom = _upload(gwn, mesh)
    for method in (gwn.BuildMethod.LBVH, gwn.BuildMethod.HPLOC):
        bvh = _build(gwn, geom, method)
        out = torch.empty(8 * n, dtype=torch.int32, device=dev)
        gwn.gwn_compute_ray_first_hit_batch(bvh, *d, out, config=gwn.QueryConfig(stack_capacity=STACK)).raise_if_error()
        torch.cuda.synchronize()
        got = out.cpu().numpy().view(gwn.RAY_HIT_DTYPE)
        ghit = got["status"] == gwn.RayStatus.HIT
        assert np.array_equal(ghit, hit), (method, np.nonzero(ghit != hit))
        tol = 5e-5 * np.maximum(1.0, np.abs(ref_t[hit]))
        assert np.all(np.abs(got["t"][hit].astype(np.float64) - ref_t[hit]) <= tol), method


# ---- dynamic contract ----------------------------------------------------------------------------

def deform(mesh):
    """tests/integration_gwn_dynamic.cu:66-77 (float arithmetic)."""
    from smallgwn_amd import synth

    f = np.float32
    x, y, z = mesh.vx, mesh.vy, mesh.vz
    nx = f(1.15) * x + f(0.10) * y + f(0.25)
    ny = f(0.85) * y - f(0.15) * z - f(0.20)
    nz = f(1.05) * z + f(0.12) * x + f(0.30)
    return synth._mesh_from_arrays(np.stack([nx, ny, nz], 1), np.stack([mesh.i0, mesh.i1, mesh.i2], 1))


def dynamic_queries(mesh):
    """tests/integration_gwn_dynamic.cu:79-139: up to 12 centroid-normal queries (point at the
    mesh diagonal, ray origin at the longer edge length, direction against the normal)."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.int64)
    span = v.max(0) - v.min(0)
    diag = np.float32(np.sqrt((span * span).sum(dtype=np.float32)))
    want = min(f.shape[0], 12)
    pts, ro, dirs = [], [], []
    for t in range(f.shape[0]):
        a, b, c = v[f[t, 0]], v[f[t, 1]], v[f[t, 2]]
        e1, e2 = b - a, c - a
        nrm = np.cross(e1, e2).astype(np.float32)
        nl = np.float32(np.sqrt((nrm * nrm).sum(dtype=np.float32)))
        if not nl > 0:
            continue
        off = np.float32(np.sqrt(max((e1 * e1).sum(dtype=np.float32), (e2 * e2).sum(dtype=np.float32))))
        cen = (a + b + c) / np.float32(3)
        un = nrm / nl
        pts.append(cen + diag * un)
        ro.append(cen + off * un)
        dirs.append(-un)
        if len(pts) == want:
            break
    P, R, D = (np.stack(x).astype(np.float32) for x in (pts, ro, dirs))
    return P, R, D


def _dynamic_run(gwn, dev, geom, bvh, mom, chain, P, R, D):
    import torch

    n = P.shape[0]
    p = _dev_arrays(dev, P[:, 0], P[:, 1], P[:, 2])
    rays = _dev_arrays(dev, R[:, 0], R[:, 1], R[:, 2], D[:, 0], D[:, 1], D[:, 2])
    cfg = gwn.QueryConfig(stack_capacity=STACK)

    def z():
        return torch.empty(n, dtype=torch.float32, device=dev)

    dist, wind, anti = z(), z(), z()
    grad, agrad = [z(), z(), z()], [z(), z(), z()]
    hits = torch.empty(8 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_unsigned_distance_batch(bvh, *p, dist, config=cfg).raise_if_error()
    gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *p, wind, config=cfg).raise_if_error()
    gwn.gwn_compute_winding_gradient_taylor_batch(bvh, mom, *p, *grad, config=cfg).raise_if_error()
    gwn.gwn_compute_winding_number_antipodal_batch(geom, bvh, chain, *p, anti, config=cfg).raise_if_error()
    gwn.gwn_compute_winding_gradient_antipodal_batch(geom, chain, *p, *agrad, config=cfg).raise_if_error()
    gwn.gwn_compute_ray_first_hit_batch(bvh, *rays, hits, config=cfg).raise_if_error()
    torch.cuda.synchronize()
    return {
        "distance": dist.cpu().numpy(), "winding": wind.cpu().numpy(), "antipodal": anti.cpu().numpy(),
        "gradient": np.stack([g.cpu().numpy() for g in grad]),
        "antipodal_gradient": np.stack([g.cpu().numpy() for g in agrad]),
        "hit": hits.cpu().numpy().view(gwn.RAY_HIT_DTYPE),
    }


def _objects(gwn, mesh, method):
    geom = _upload(gwn, mesh)
    bvh = _build(gwn, geom, method)
    mom = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain).raise_if_error()
    return geom, bvh, mom, chain


def dynamic_contract(gwn, dev, mesh) -> None:
    moved = deform(mesh)
    P, R, D = dynamic_queries(moved)
    assert P.shape[0] > 0, "no non-degenerate triangle"
    for method in (gwn.BuildMethod.LBVH, gwn.BuildMethod.HPLOC):
        geom, bvh, mom, chain = _objects(gwn, mesh, method)
        initial = _dynamic_run(gwn, dev, geom, bvh, mom, chain, P, R, D)
        gwn.gwn_update_geometry(geom, moved.vx, moved.vy, moved.vz).raise_if_error()
        gwn.gwn_refit_bvh(geom, bvh).raise_if_error()
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        refit = _dynamic_run(gwn, dev, geom, bvh, mom, chain, P, R, D)
        fresh = _dynamic_run(gwn, dev, *_objects(gwn, moved, method), P, R, D)
        assert np.any(np.abs(refit["distance"] - initial["distance"]) > 1e-4), "deformation changed nothing"
        assert np.all(np.abs(refit["distance"] - fresh["distance"]) <= 2e-5)
        assert np.all(np.abs(refit["winding"] - fresh["winding"]) <= 2e-3)
        assert np.all(np.abs(refit["antipodal"] - fresh["antipodal"]) <= 2e-5)
        assert np.all(np.abs(refit["gradient"] - fresh["gradient"]) <= 2e-3)
        assert np.all(np.abs(refit["antipodal_gradient"] - fresh["antipodal_gradient"]) <= 2e-5)
        rh, fh = refit["hit"], fresh["hit"]
        assert np.array_equal(rh["status"], fh["status"])
        assert np.all(rh["status"] == gwn.RayStatus.HIT)
        assert np.all(np.abs(rh["t"] - fh["t"]) <= 2e-5)
        same = rh["primitive_id"] == fh["primitive_id"]
        for k in ("u", "v", "nx", "ny", "nz"):
            assert np.all(np.abs(rh[k][same] - fh[k][same]) <= 2e-5), k
        v = np.stack([moved.vx, moved.vy, moved.vz], 1)
        f = np.stack([moved.i0, moved.i1, moved.i2], 1).astype(np.int64)
        for h in (rh, fh):
            pid = h["primitive_id"].astype(np.int64)
            assert (pid < f.shape[0]).all()
            a, b, c = v[f[pid, 0]], v[f[pid, 1]], v[f[pid, 2]]
            expect = np.cross(b - a, c - a)
            got = np.stack([h["nx"], h["ny"], h["nz"]], 1)
            assert np.all((expect * got).sum(1) > 0)


def run_directory(gwn, orc, dev, mesh_dir: str, log=print) -> int:
    """Every contract on every loadable mesh of a directory; returns the number of meshes tested."""
    from smallgwn_amd import ply

    tested = 0
    for path in ply.collect_ply_model_paths(mesh_dir):
        mesh = ply.load_ply_mesh(path)
        if mesh is None:
            log(f"smallgwn input skip: failed to parse {path}")
            continue
        log(f"[mesh] {path}: {mesh.vertex_count} vertices, {mesh.triangle_count} triangles")
        builder_contract(gwn, orc, dev, mesh)
        static_contract(gwn, dev, mesh)
        dynamic_contract(gwn, dev, mesh)
        tested += 1
    return tested


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--mesh-dir", required=True)
    args = ap.parse_args()
    import torch

    from oracle import oracle
    from smallgwn_amd import gwn

    oracle.lib()
    n = run_directory(gwn, oracle, torch.device("cuda:0"), args.mesh_dir)
    print(f"{n} meshes passed")
    return 0 if n > 0 else 1


if __name__ == "__main__":
    sys.exit(main())
