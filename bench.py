"""Throughput of the geometric-query hot path on one MI355X.

Workload (large single mesh): radially displaced sphere r = 1 + 0.05 * (8 random-phase real
spherical harmonics of degree <= 12), 1024 x 512 latitude-longitude quads = 1,048,576
triangles; 16,777,216 query points, half uniform in [-1.5, 1.5]^3 and half within 0.01 of the
surface along the face normal.  Seeded synthetic data (splitmix64 from 0xB3A9E5D4).

One step = one Taylor order-2 winding batch over all 16M queries (Hilbert presort + traversal),
FP32 traversal and far field, near-field triangle terms with FP64 numerators and denominators,
at an accuracy scale (6.25) that keeps max |dW| <= 1e-4 against the FP64 exact winding number on a
seeded, strided subsample of 1,048,576 queries.  Inputs, BVH and moments are resident in HBM
before the timed region.

Also reported (outside the timed steps), each with its roofline where the kernel has one:
- the exact winding batch with the same FP64 near field on 524,288 queries (enough to fill the
  GPU), the plain form the Taylor path is compared with;
- BVH build (H-PLOC radius 8, LBVH) and moment refit throughput against compulsory HBM bytes;
- the small configuration (icosphere 4x, 65,536 queries and rays);
- the large mesh's other batches (Taylor at scale 2, unsigned distance, first-hit rays);
- the precision-split report (262K-triangle open soup, 4,194,304 log-offset queries);
- the differentiation step (131,072 triangles, 1,048,576 queries, 262,144 targets);
- the batched mode (4,096 meshes of 200 to 200,000 triangles and 64 to 1,048,576 queries);
- the CPU baseline: the oracle's C replay of the same traversal on the same H-PLOC tree, timed on
  the host's cores, plus the small configuration on the CPU and the accuracy of the
  differentiation step's gradient against the FP64 oracle adjoint.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

METRIC = "Mqueries/s Taylor vs exact at |dW|<=1e-4 vs FP64; adjoint steps/s; build Mtri/s"

# Algorithmic FLOP weights per unit of traversal work (special functions count 1 each).
FLOPS_CHILD_TEST = 10
# An accepted expansion: the value-only queries evaluate the folded forms of the evaluation
# records (11 / 27 / 55 FLOP); the gradient and adjoint kernels the reference's expression on the
# traversal records (14 / 34 / 79).
FLOPS_ACCEPT = {0: 11, 1: 27, 2: 55}
FLOPS_ACCEPT_REFERENCE = {0: 14, 1: 34, 2: 79}
FLOPS_TRIANGLE = 87            # solid angle: 80 FLOP + 3 sqrt + 3 div + 1 atan2
FLOPS_DIST_BOX = 12            # point-box distance per child
FLOPS_DIST_TRI = 45            # point-triangle distance
FLOPS_RAY_SLAB = 14            # ray-box slab test per child
FLOPS_RAY_TRI = 60             # oriented Pluecker ray-triangle test
# Adjoint extras on top of the forward's work with the same decisions: 12 field adjoints of an
# accepted order-1 slot (dw/dP, dw/dN, dw/dT), and the 9 vertex adjoints of a near triangle.
FLOPS_ADJ_ACCEPT = 60
FLOPS_ADJ_TRIANGLE = 120
PEAK_FP32_TFLOPS = 157.3       # MI355X dense FP32 vector
HBM_PEAK_GBPS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--queries", type=int, default=16_777_216)
    p.add_argument("--order", type=int, default=2)
    p.add_argument("--beta", type=float, default=6.25, help="Taylor accuracy scale")
    p.add_argument("--check-queries", type=int, default=1_048_576, help="FP64 subsample for |dW|")
    p.add_argument("--exact-queries", type=int, default=524_288, help="queries timed for the exact form")
    p.add_argument("--cpu-seconds", type=float, default=3.0, help="target duration of one CPU baseline run")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiation workload")
    p.add_argument("--adjoint-steps", type=int, default=5)
    p.add_argument("--no-batched", action="store_true", help="skip the batched-mode workload")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the small configuration, the large-mesh extras and the precision-split report")
    p.add_argument("--builder", choices=("hploc", "lbvh"), default="hploc", help="BVH builder of the large mesh")
    p.add_argument("--batched-meshes", type=int, default=4096)
    p.add_argument("--batched-tune-build", action="store_true",
                   help="also time the build tuner (widths x H-PLOC radii) of the batch over geometries")
    p.add_argument("--precision-queries", type=int, default=4_194_304)
    return p.parse_args()


def _event_ms(stream, fn, reps=3) -> float:
    """Average milliseconds of fn() between HIP events on the launch stream, after one warm-up."""
    import torch

    fn()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    e[0].record(stream)
    for _ in range(reps):
        fn()
    e[1].record(stream)
    torch.cuda.synchronize()
    return e[0].elapsed_time(e[1]) / reps


def _valu_roofline(flops: float, ms: float, kernel: str, note: str) -> dict:
    tflops = flops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / PEAK_FP32_TFLOPS, 4), "kernel": kernel, "kernel_ms": round(ms, 4),
            "flops_per_launch": float(flops), "note": note}


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (queries shard by "
              "contiguous ranges with a replicated BVH; nothing here uses a second GPU)", file=sys.stderr)
        return 2

    import torch

    from smallgwn_amd import gwn, synth

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    t0 = time.time()
    mesh = synth.config_large_mesh()
    qx, qy, qz = synth.mixed_queries(mesh, args.queries, near_fraction=0.5, band=0.01)
    log(f"inputs: {mesh.triangle_count} tris, {qx.shape[0]} queries ({time.time() - t0:.1f}s)")

    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    if args.builder == "hploc":
        opts = gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=8)
    else:
        opts = gwn.BuildOptions(method=gwn.BuildMethod.LBVH)
    gwn.gwn_build_bvh(geom, bvh, opts).raise_if_error()
    mom = gwn.MomentObject(args.order)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    info = bvh.info()
    d = [torch.from_numpy(a).to(dev) for a in (qx, qy, qz)]
    n = d[0].numel()
    out = torch.empty(n, dtype=torch.float32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    cfg = gwn.QueryConfig(stack_capacity=max(64, info["internal_stack_bound"]))
    # A plain 256 MiB device copy of known size, outside every timed region: under rocprofv3
    # --pmc it calibrates FETCH_SIZE / WRITE_SIZE against bytes actually moved.
    calib_src = torch.ones(1 << 26, dtype=torch.float32, device=dev)
    calib_dst = torch.empty_like(calib_src)
    calib_dst.copy_(calib_src)
    del calib_src, calib_dst
    torch.cuda.synchronize()

    def step():
        gwn.gwn_compute_winding_number_taylor_batch(
            bvh, mom, *d, out, accuracy_scale=args.beta, config=cfg, precision=gwn.Precision.SPLIT,
        ).raise_if_error()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(args.steps):
        step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms_per_step = ev[0].elapsed_time(ev[1]) / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    log(f"taylor o{args.order} split beta={args.beta}: {ms_per_step:.3f} ms/step, {value:.1f} Mq/s")

    # Stage split: presort vs traversal kernel, timed with events on the launch stream.
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = max(3, args.steps)
    t_order = t_trav = 0.0
    for _ in range(reps):
        e[0].record(stream)
        gwn.gwn_compute_query_order(bvh, *d, order).raise_if_error()
        e[1].record(stream)
        gwn.gwn_compute_winding_number_taylor_batch(
            bvh, mom, *d, out, accuracy_scale=args.beta, config=cfg, precision=gwn.Precision.SPLIT, order=order,
        ).raise_if_error()
        e[2].record(stream)
        torch.cuda.synchronize()
        t_order += e[0].elapsed_time(e[1]) / reps
        t_trav += e[1].elapsed_time(e[2]) / reps

    # Work counters of the traversal (one untimed pass) -> algorithmic FLOPs.
    cnt = torch.zeros(3 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(
        bvh, mom, *d, out, accuracy_scale=args.beta, config=cfg, precision=gwn.Precision.SPLIT,
        order=order, counters=cnt,
    ).raise_if_error()
    c = cnt.view(n, 3).to(torch.int64).sum(0).cpu().numpy()
    tests, accepts, leaves = (int(v) for v in c)
    del cnt
    # Every near-field triangle at its FP32 weight: its FP64 part (numerator and denominator in the
    # near zone) is not weighted up, so the figure is a lower bound on the FLOPs done.
    fp32_equiv = tests * FLOPS_CHILD_TEST + accepts * FLOPS_ACCEPT[args.order] + leaves * FLOPS_TRIANGLE
    achieved_tflops = fp32_equiv / (t_trav * 1e-3) / 1e12

    # Accuracy against FP64 exact on a seeded, strided subsample (both query halves).
    taylor_host = out.cpu().numpy()
    m = min(args.check_queries, n)
    idx = np.unique(np.linspace(0, n - 1, m).astype(np.int64))
    sub = [t[torch.from_numpy(idx).to(dev)].contiguous() for t in d]
    ref = torch.empty(idx.size, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *sub, ref, gwn.Precision.FP64).raise_if_error()
    torch.cuda.synchronize()
    err = np.abs(taylor_host[idx].astype(np.float64) - ref.cpu().numpy())
    max_err, p99_err = float(err.max()), float(np.quantile(err, 0.99))
    log(f"accuracy on {idx.size} queries: max |dW| {max_err:.3e}, p99 {p99_err:.3e}")

    # The exact form with the same FP64 near field, on enough queries to fill the GPU (cost is
    # linear in the queries once every SIMD holds several waves).
    me = min(args.exact_queries, n)
    eidx = torch.from_numpy(np.linspace(0, n - 1, me).astype(np.int64)).to(dev)
    esub = [t[eidx].contiguous() for t in d]
    eout = torch.empty(me, dtype=torch.float64, device=dev)
    exact_ms = _event_ms(stream, lambda: gwn.gwn_compute_winding_number_exact_batch_ex(
        bvh, *esub, eout, gwn.Precision.SPLIT).raise_if_error(), reps=1)
    exact_rate = me / (exact_ms * 1e-3) / 1e6
    exact_flops = float(me) * mesh.triangle_count * FLOPS_TRIANGLE
    log(f"exact split: {exact_rate:.4f} Mq/s on {me} queries ({exact_ms:.1f} ms)")
    del esub, eout

    # Build and refit throughput.
    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    # Both builders on their own objects (the large mesh's tree is rebuilt deterministically).
    lbvh_bvh, hploc_bvh = gwn.BvhObject(), gwn.BvhObject()
    lbvh_opts = gwn.BuildOptions(method=gwn.BuildMethod.LBVH)
    hploc_opts = gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=8)
    lbvh_s = timed(lambda: gwn.gwn_build_bvh(geom, lbvh_bvh, lbvh_opts).raise_if_error(), 5)
    hploc_s = timed(lambda: gwn.gwn_build_bvh(geom, hploc_bvh, hploc_opts).raise_if_error(), 5)
    lbvh_info, hploc_info = lbvh_bvh.info(), hploc_bvh.info()
    del lbvh_bvh, hploc_bvh
    build_s = hploc_s if args.builder == "hploc" else lbvh_s
    mom_t = gwn.MomentObject(args.order)
    refit_s = timed(lambda: gwn.gwn_refit_bvh_moment(bvh, mom_t).raise_if_error(), 5)
    del mom_t
    bounds_s = timed(lambda: gwn.gwn_refit_bvh(geom, bvh).raise_if_error(), 5)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()  # moments of the refit (unchanged) bounds
    T = mesh.triangle_count
    log(f"build H-PLOC r8 {hploc_s * 1e3:.2f} ms ({T / hploc_s / 1e6:.1f} Mtri/s), LBVH {lbvh_s * 1e3:.2f} ms "
        f"({T / lbvh_s / 1e6:.1f} Mtri/s), refit o{args.order} {refit_s * 1e3:.2f} ms, bounds refit "
        f"{bounds_s * 1e3:.2f} ms")

    traffic = traffic_profile()
    valu = valu_profile()
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "taylor_winding_large_mesh",
            "triangles": T,
            "queries": n,
            "near_surface_fraction": 0.5,
            "taylor_order": args.order,
            "accuracy_scale": args.beta,
            "precision": "fp32 traversal/far field, fp64 near-field numerator/denominator",
            "builder": "hploc_r8" if args.builder == "hploc" else "lbvh",
            "bvh_width": 4,
        },
        "accuracy": {"max_abs_dw": max_err, "p99_abs_dw": p99_err, "checked_queries": int(idx.size),
                     "target": 1e-4, "meets_target": bool(max_err <= 1e-4)},
        "exact_fp64_near_field": {
            "mqueries_per_s": exact_rate, "timed_queries": me, "ms": exact_ms,
            "taylor_speedup": value / exact_rate,
            "roofline": _valu_roofline(exact_flops, exact_ms, "exact_kernel<split>",
                                       "query x triangle pairs at the FP32 solid-angle weight (87); the "
                                       "near-zone FP64 part is not weighted up")},
        "stages_ms": {"query_presort": round(t_order, 4), "traversal": round(t_trav, 4)},
        "build": {"method": "hploc_r8" if args.builder == "hploc" else "lbvh", "ms": build_s * 1e3,
                  "mtri_per_s": T / build_s / 1e6,
                  "refit_moment_ms": refit_s * 1e3, "refit_mtri_per_s": T / refit_s / 1e6,
                  "refit_bounds_ms": bounds_s * 1e3,
                  "hploc_r8_ms": hploc_s * 1e3, "hploc_r8_mtri_per_s": T / hploc_s / 1e6,
                  "hploc_nodes": hploc_info["node_count"],
                  "hploc_internal_stack_bound": hploc_info["internal_stack_bound"],
                  "lbvh_ms": lbvh_s * 1e3, "lbvh_mtri_per_s": T / lbvh_s / 1e6,
                  "lbvh_nodes": lbvh_info["node_count"],
                  "lbvh_internal_stack_bound": lbvh_info["internal_stack_bound"],
                  "roofline": build_roofline(T, hploc_s, lbvh_s, refit_s, bounds_s, info["node_count"],
                                             args.order)},
        "work": {"child_tests": tests, "accepted": accepts, "near_field_triangles": leaves},
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved_tflops, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic.get("hbm_bytes_per_launch"),
            "traffic_source": traffic.get("source"),
            "traffic_current": traffic.get("current"),
            "kernel": f"taylor_wave_kernel<{args.order}, split>",
            "kernel_ms": round(t_trav, 4),
            "flops_per_launch": fp32_equiv,
            "valu_lane_utilization": valu.get("valu_lane_utilization"),
            "valu_source": valu.get("source"),
            "valu_current": valu.get("current"),
            "note": "FP32 FLOPs from per-query work counters (10 per child test, 55 per accepted "
                    "order-2 expansion in its folded form, 87 per near-field triangle); peak is MI355X's FP32 vector "
                    "figure, which counts packed FP32, disabled on this platform (non-packed 78.6); "
                    "traffic and lane utilization come from the rocprofv3 --pmc summaries named in "
                    "traffic_source / valu_source, used only while the kernel's device-source digest "
                    "equals the one they were measured at (traffic_current / valu_current)",
        },
    }
    del sub, ref

    if not args.no_extras:
        result["config_small"] = small_config(dev, stream)
        result["large_mesh_extras"] = large_extras(args, dev, stream, bvh, d, idx)
        result["precision_split"] = precision_report(args, dev, stream)

    del d, out, order
    adjoint_check = None
    if not args.no_adjoint:
        result["adjoint"], adjoint_check = adjoint_bench(args, dev, stream)

    if not args.no_batched:
        result["batched"] = batched_bench(args, dev, stream)

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(mesh, qx, qy, qz, args, value)
        if not args.no_extras:
            result["config_small"]["cpu"] = cpu_small()
        if adjoint_check is not None:
            result["adjoint"]["gradient_check"] = adjoint_gradient_check(adjoint_check)
            result["adjoint"]["gradient_rel_error"] = result["adjoint"]["gradient_check"]["rel_error"]

    print(json.dumps(result), flush=True)
    return 0


# Compulsory HBM bytes of a build: about 800 per triangle (keys, ids and boxes through the sort
# passes, binary nodes, the collapse and the records; LBVH with u64 keys).  A moment refit reads
# each wide node (128 B) and its triangles (36 B each), and writes and reads the node's payload
# (136 B) and writes its record (order 0/1/2: 128/256/384 B).  A bounds refit reads each triangle's
# indices and vertices (12 + 36 B) and writes its record (36 B), and reads and writes each node
# (128 + 128 B) and its 24-byte box.
BUILD_BYTES_PER_TRIANGLE = 800


def build_roofline(tris: int, hploc_s: float, lbvh_s: float, refit_s: float, bounds_s: float, nodes: int,
                   order: int) -> dict:
    """HBM roofline of the builds and refits against their compulsory bytes."""
    record = {0: 128, 1: 256, 2: 384}[order]
    refit_bytes = nodes * (128 + 136 + 136 + record) + tris * 36
    bounds_bytes = tris * (12 + 36 + 36) + nodes * (128 + 128 + 24 + 24)
    build_bytes = BUILD_BYTES_PER_TRIANGLE * tris
    out = {"bound": "hbm", "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "bytes_per_triangle_build": BUILD_BYTES_PER_TRIANGLE, "refit_bytes": refit_bytes,
           "refit_bounds_bytes": bounds_bytes}
    for name, nbytes, sec in (("hploc_r8", build_bytes, hploc_s), ("lbvh", build_bytes, lbvh_s),
                              ("refit_moment", refit_bytes, refit_s), ("refit_bounds", bounds_bytes, bounds_s)):
        gbps = nbytes / sec / 1e9
        out[name] = {"achieved": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}
    return out


def small_config(dev, stream) -> dict:
    """The small configuration on the GPU: icosphere 4x (5,120 triangles), 65,536 queries in
    [-1.5, 1.5]^3 and 65,536 rays, LBVH, BVH4, Taylor order 1 at scale 2; exact and Taylor winding,
    unsigned distance and first-hit rays."""
    import torch

    from smallgwn_amd import gwn, synth

    mesh, q, rays = synth.config_small()
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, mesh.vx, mesh.vy, mesh.vz, mesh.i0, mesh.i1, mesh.i2).raise_if_error()
    bvh = gwn.BvhObject()
    lbvh = gwn.BuildOptions(method=gwn.BuildMethod.LBVH)
    gwn.gwn_build_bvh(geom, bvh, lbvh).raise_if_error()
    mom = gwn.MomentObject(1)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    d = [torch.from_numpy(a).to(dev) for a in q]
    r = [torch.from_numpy(a).to(dev) for a in rays]
    n = d[0].numel()
    out = torch.empty(n, dtype=torch.float32, device=dev)
    hits = torch.empty(8 * n, dtype=torch.int32, device=dev)
    t = time.perf_counter()
    for _ in range(10):
        gwn.gwn_build_bvh(geom, bvh, lbvh).raise_if_error()
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    build_ms = (time.perf_counter() - t) / 10 * 1e3
    res = {"triangles": mesh.triangle_count, "queries": n, "rays": n, "builder": "lbvh", "taylor_order": 1,
           "accuracy_scale": 2.0, "build_and_refit_ms": build_ms}
    for name, fn in (
            ("exact", lambda: gwn.gwn_compute_winding_number_exact_batch(bvh, *d, out)),
            ("taylor_o1", lambda: gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, *d, out, accuracy_scale=2.0)),
            ("distance", lambda: gwn.gwn_compute_unsigned_distance_batch(bvh, *d, out)),
            ("ray_first_hit", lambda: gwn.gwn_compute_ray_first_hit_batch(bvh, *r, hits))):
        ms = _event_ms(stream, lambda: fn().raise_if_error(), reps=5)
        res[name] = {"ms": ms, "mqueries_per_s": n / (ms * 1e-3) / 1e6}
    log(f"small config: " + ", ".join(f"{k} {v['mqueries_per_s']:.0f} Mq/s" for k, v in res.items()
                                      if isinstance(v, dict)))
    return res


def _distance_ray_work(dev, bvh, d, rays, gwn):
    import torch

    n = d[0].numel()
    work = torch.zeros(2 * n, dtype=torch.int32, device=dev)
    gwn.gwn_compute_unsigned_distance_work(bvh, *d, work).raise_if_error()
    dw = work.view(n, 2).to(torch.int64).sum(0).cpu().numpy()
    gwn.gwn_compute_ray_first_hit_work(bvh, *rays, work).raise_if_error()
    rw = work.view(n, 2).to(torch.int64).sum(0).cpu().numpy()
    return (int(dw[0]), int(dw[1])), (int(rw[0]), int(rw[1]))


def large_extras(args, dev, stream, bvh, d, idx) -> dict:
    """The large mesh's other batches on the same tree and queries: Taylor orders 1 and 2 at the
    reference's default accuracy scale 2 (plain precision, with their error on the checked
    subsample), the unsigned distance of all 16M queries, and 16,777,216 first-hit rays (origins on a
    radius-3 sphere, directions toward uniform points in the unit ball), with the distance and ray
    rooflines from their traversals' work counters."""
    import torch

    from smallgwn_amd import gwn, synth

    n = d[0].numel()
    out = torch.empty(n, dtype=torch.float32, device=dev)
    sub = [t[torch.from_numpy(idx).to(dev)].contiguous() for t in d]
    ref = torch.empty(idx.size, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *sub, ref, gwn.Precision.FP64).raise_if_error()
    ref_h = ref.cpu().numpy()
    del sub, ref
    res = {}
    for order in (1, 2):
        mom = gwn.MomentObject(order)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        ms = _event_ms(stream, lambda: gwn.gwn_compute_winding_number_taylor_batch(
            bvh, mom, *d, out, accuracy_scale=2.0).raise_if_error())
        err = np.abs(out.cpu().numpy()[idx].astype(np.float64) - ref_h)
        res[f"taylor_o{order}_scale2"] = {"mqueries_per_s": n / (ms * 1e-3) / 1e6, "ms": ms,
                                          "max_abs_dw": float(err.max()),
                                          "p99_abs_dw": float(np.quantile(err, 0.99))}
        log(f"taylor o{order} plain scale 2: {ms:.2f} ms ({n / (ms * 1e-3) / 1e6:.1f} Mq/s), "
            f"max |dW| {err.max():.3e}")
        del mom
    rays = [torch.from_numpy(a).to(dev) for a in synth.rays_toward_ball(n, seed=synth.DEFAULT_SEED + 3)]
    (dbox, dtri), (rbox, rtri) = _distance_ray_work(dev, bvh, d, rays, gwn)
    dist_flops = dbox * FLOPS_DIST_BOX + dtri * FLOPS_DIST_TRI
    ray_flops = rbox * FLOPS_RAY_SLAB + rtri * FLOPS_RAY_TRI
    ms = _event_ms(stream, lambda: gwn.gwn_compute_unsigned_distance_batch(bvh, *d, out).raise_if_error())
    res["distance"] = {"mqueries_per_s": n / (ms * 1e-3) / 1e6, "ms": ms, "queries": n,
                       "traversal": "reference order, four lanes per point (bit-exact)",
                       "work": {"box_tests": dbox, "triangle_tests": dtri},
                       "roofline": _valu_roofline(dist_flops, ms, "distance_quad_kernel",
                                                  "12 FLOP per point-box test, 45 per point-triangle distance, "
                                                  "counted by the one-lane traversal the quad replays")}
    log(f"distance: {ms:.2f} ms ({n / (ms * 1e-3) / 1e6:.1f} Mq/s), frac {res['distance']['roofline']['frac']}")
    lane_cfg = gwn.QueryConfig(lanes_per_point=1)
    ms = _event_ms(stream, lambda: gwn.gwn_compute_unsigned_distance_batch(bvh, *d, out, config=lane_cfg).raise_if_error())
    res["distance_one_lane"] = {"mqueries_per_s": n / (ms * 1e-3) / 1e6, "ms": ms, "queries": n,
                                "traversal": "reference order, one lane per point (bit-exact)"}
    wave_cfg = gwn.QueryConfig(sort_queries=2)
    ms = _event_ms(stream, lambda: gwn.gwn_compute_unsigned_distance_batch(bvh, *d, out, config=wave_cfg).raise_if_error())
    res["distance_wave"] = {"mqueries_per_s": n / (ms * 1e-3) / 1e6, "ms": ms, "queries": n,
                            "traversal": "wave-cooperative (sort_queries=2, within FP32 ulps)"}
    del out
    hits = torch.empty(8 * n, dtype=torch.int32, device=dev)
    ms = _event_ms(stream, lambda: gwn.gwn_compute_ray_first_hit_batch(bvh, *rays, hits).raise_if_error())
    status = hits.view(torch.uint8).view(n, 32)[:, 28]
    hit_frac = float((status == int(gwn.RayStatus.HIT)).to(torch.float32).mean())
    res["ray_first_hit"] = {"mrays_per_s": n / (ms * 1e-3) / 1e6, "ms": ms, "rays": n, "hit_fraction": hit_frac,
                            "work": {"slab_tests": rbox, "triangle_tests": rtri},
                            "roofline": _valu_roofline(ray_flops, ms, "ray_kernel<packed>",
                                                       "14 FLOP per slab test, 60 per Pluecker triangle test")}
    log(f"rays: {ms:.2f} ms ({n / (ms * 1e-3) / 1e6:.1f} Mrays/s), hit fraction {hit_frac:.3f}, "
        f"frac {res['ray_first_hit']['roofline']['frac']}")
    return res


def precision_report(args, dev, stream) -> dict:
    """Precision-split configuration: open soup of ~262K triangles (displaced sphere, 30% of faces
    deleted, 2% with jittered duplicate vertices), 4,194,304 queries at normal offsets log-uniform in
    [1e-6, 1e-1] from random surface points.  Taylor order 2 (at the headline accuracy scale) runs on
    all queries; the exact and antipodal batches, plain (FP32) and split (FP64 near field), on a
    strided 262,144-query subsample; max and p99 |dW| against the all-FP64 exact form on that
    subsample, and throughput per mode."""
    import torch

    from smallgwn_amd import gwn, synth

    t0 = time.time()
    soup = synth.open_soup(612, 306, seed=synth.DEFAULT_SEED)
    q = synth.near_surface_log_queries(soup, args.precision_queries, seed=synth.DEFAULT_SEED + 5, lo=1e-6, hi=1e-1)
    n = q[0].shape[0]
    geom = gwn.GeometryObject()
    gwn.gwn_upload_geometry(geom, soup.vx, soup.vy, soup.vz, soup.i0, soup.i1, soup.i2).raise_if_error()
    bvh = gwn.BvhObject()
    gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC, hploc_search_radius=8)).raise_if_error()
    mom = gwn.MomentObject(2)
    gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
    chain = gwn.BoundaryChainObject()
    gwn.gwn_build_boundary_chain(geom, chain).raise_if_error()
    d = [torch.from_numpy(a).to(dev) for a in q]
    m = min(262_144, n)
    idx = np.linspace(0, n - 1, m).astype(np.int64)
    didx = torch.from_numpy(idx).to(dev)
    sub = [a[didx].contiguous() for a in d]
    ref = torch.empty(m, dtype=torch.float64, device=dev)
    gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *sub, ref, gwn.Precision.FP64).raise_if_error()
    ref_h = ref.cpu().numpy()
    full = torch.empty(n, dtype=torch.float32, device=dev)
    o32 = torch.empty(m, dtype=torch.float32, device=dev)
    o64 = torch.empty(m, dtype=torch.float64, device=dev)
    stack = gwn.QueryConfig(stack_capacity=max(64, bvh.info()["internal_stack_bound"]))
    modes = {
        "exact_plain": (o32, m, lambda: gwn.gwn_compute_winding_number_exact_batch(bvh, *sub, o32)),
        "exact_split": (o64, m, lambda: gwn.gwn_compute_winding_number_exact_batch_ex(bvh, *sub, o64, gwn.Precision.SPLIT)),
        "taylor_o2_plain": (full, n, lambda: gwn.gwn_compute_winding_number_taylor_batch(
            bvh, mom, *d, full, accuracy_scale=args.beta, config=stack)),
        "taylor_o2_split": (full, n, lambda: gwn.gwn_compute_winding_number_taylor_batch(
            bvh, mom, *d, full, accuracy_scale=args.beta, precision=gwn.Precision.SPLIT, config=stack)),
        "antipodal_plain": (o32, m, lambda: gwn.gwn_compute_winding_number_antipodal_batch(
            geom, bvh, chain, *sub, o32, config=stack)),
        "antipodal_split": (o32, m, lambda: gwn.gwn_compute_winding_number_antipodal_batch(
            geom, bvh, chain, *sub, o32, precision=gwn.Precision.SPLIT, config=stack)),
    }
    res = {"triangles": soup.triangle_count, "queries": n, "checked_queries": m, "accuracy_scale": args.beta,
           "boundary_edges": chain.info()["edge_count"]}
    for name, (buf, count, fn) in modes.items():
        ms = _event_ms(stream, lambda: fn().raise_if_error(), reps=1)
        w = buf.cpu().numpy().astype(np.float64)
        if count == n:
            w = w[idx]
        ok = np.isfinite(w)
        err = np.abs(w[ok] - ref_h[ok])
        res[name] = {"max_abs_dw": float(err.max()), "p99_abs_dw": float(np.quantile(err, 0.99)),
                     "nan_results": int((~ok).sum()), "timed_queries": count,
                     "mqueries_per_s": count / (ms * 1e-3) / 1e6}
        log(f"precision {name}: max |dW| {err.max():.3e}, p99 {np.quantile(err, 0.99):.3e}, "
            f"{count / (ms * 1e-3) / 1e6:.2f} Mq/s on {count}")
    log(f"precision report: {soup.triangle_count} tris, {n} queries ({time.time() - t0:.1f}s)")
    return res


def adjoint_bench(args, dev, stream):
    """Differentiation workload: 131,072-triangle displaced sphere, 1,048,576 labelled queries
    (inside a radius-0.9 sphere or not), 262,144 target points on that sphere.  One step: new
    vertex positions (a gradient step), BVH bounds and order-1 moment refit, Taylor winding and
    distance forward, loss, adjoint backward to dL/dV.  Also returns the inputs of the gradient
    check (the winding adjoint on a strided query subsample), which the CPU leg compares with the
    FP64 oracle adjoint."""
    import torch

    from smallgwn_amd import autograd as ag
    from smallgwn_amd import gwn, synth

    mesh = synth.displaced_sphere(256, 256)
    faces = np.stack([mesh.i0, mesh.i1, mesh.i2], 1)
    initial = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    verts = torch.tensor(initial, device=dev, requires_grad=True)
    scene = ag.MeshScene(faces, verts, order=1, method=gwn.BuildMethod.HPLOC, accuracy_scale=2.0)
    qx, qy, qz = synth.uniform_box_points(21, 1 << 20, (-1.5,) * 3, (1.5,) * 3)
    queries = torch.tensor(np.stack([qx, qy, qz], 1), device=dev)
    labels = (queries.norm(dim=1) < 0.9).to(torch.float32)
    tx, ty, tz = synth.uniform_box_points(22, 1 << 18, (-1.0,) * 3, (1.0,) * 3)
    targets = torch.tensor(np.stack([tx, ty, tz], 1), device=dev)
    targets = 0.9 * targets / targets.norm(dim=1, keepdim=True)
    opt = torch.optim.SGD([verts], lr=1e-3)

    def step():
        opt.zero_grad(set_to_none=True)
        w = ag.taylor_winding(verts, queries, scene)
        dist = ag.unsigned_distance(verts, targets, scene)
        loss = ((torch.sigmoid(20.0 * (w - 0.5)) - labels) ** 2).mean() + dist.mean()
        loss.backward()
        opt.step()
        return loss

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(args.adjoint_steps):
        loss = step()
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / args.adjoint_steps
    log(f"adjoint: {ms:.2f} ms/step ({1e3 / ms:.1f} steps/s), loss {float(loss):.5f}")

    # The winding adjoint kernel alone on the current positions, with its forward's work counters
    # (the adjoint takes the same accept/descend decisions) for its roofline.
    scene.update(verts)
    V = mesh.vertex_count
    q = [queries[:, k].contiguous() for k in range(3)]
    w = torch.empty(queries.shape[0], dtype=torch.float32, device=dev)
    nq = queries.shape[0]
    cnt = torch.zeros(3 * nq, dtype=torch.int32, device=dev)
    gwn.gwn_compute_winding_number_taylor_batch(scene.bvh, scene.moment, *q, w, accuracy_scale=2.0,
                                                counters=cnt).raise_if_error()
    tests, accepts, leaves = (int(v) for v in cnt.view(nq, 3).to(torch.int64).sum(0).cpu().numpy())
    sig = torch.sigmoid(20.0 * (w - 0.5))
    u = (2.0 * (sig - labels) * sig * (1.0 - sig) * 20.0 / nq).contiguous()
    dv = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]

    def adj():
        gwn.gwn_compute_winding_taylor_adjoint_batch(scene.geometry, scene.bvh, scene.moment, *q, u, *dv,
                                                     accuracy_scale=2.0).raise_if_error()

    adj_ms = _event_ms(stream, adj, reps=3)
    fwd_ms = _event_ms(stream, lambda: gwn.gwn_compute_winding_number_taylor_batch(
        scene.bvh, scene.moment, *q, w, accuracy_scale=2.0).raise_if_error(), reps=3)
    adj_flops = (tests * FLOPS_CHILD_TEST + accepts * (FLOPS_ACCEPT_REFERENCE[1] + FLOPS_ADJ_ACCEPT)
                 + leaves * (FLOPS_TRIANGLE + FLOPS_ADJ_TRIANGLE))
    res = {"steps_per_s": 1e3 / ms, "ms_per_step": ms, "triangles": mesh.triangle_count,
           "queries": nq, "targets": int(targets.shape[0]), "taylor_order": 1,
           "builder": "hploc_r8", "step": "update + bounds/moment refit + winding and distance forward "
                                          "+ adjoint backward + SGD update",
           "winding_forward_ms": fwd_ms, "winding_adjoint_ms": adj_ms,
           "work": {"child_tests": tests, "accepted": accepts, "near_field_triangles": leaves},
           "roofline": _valu_roofline(adj_flops, adj_ms, "winding adjoint batch (Hilbert order, wave traversal)",
                                      "forward weights (10 / 34 / 87) plus 60 FLOP per accepted slot's 12 field "
                                      "adjoints and 120 per near triangle's 9 vertex adjoints; work counted by "
                                      "the forward with the same decisions; includes the Hilbert presort")}
    log(f"adjoint kernel {adj_ms:.2f} ms (forward {fwd_ms:.2f} ms), frac {res['roofline']['frac']}")
    # Gradient check inputs: the adjoint on 4,096 strided queries with the step's upstream gradient.
    sidx = np.linspace(0, nq - 1, 4096).astype(np.int64)
    ts = torch.from_numpy(sidx).to(dev)
    gs = [torch.zeros(V, dtype=torch.float32, device=dev) for _ in range(3)]
    gwn.gwn_compute_winding_taylor_adjoint_batch(scene.geometry, scene.bvh, scene.moment,
                                                 *[a[ts].contiguous() for a in q], u[ts].contiguous(), *gs,
                                                 accuracy_scale=2.0).raise_if_error()
    torch.cuda.synchronize()
    vpos = verts.detach().cpu().numpy().astype(np.float32)
    check = {"faces": faces, "initial": initial, "vertices": vpos, "queries": [np.ascontiguousarray(a[ts].cpu().numpy()) for a in q],
             "u": u[ts].double().cpu().numpy(), "gpu": np.stack([g.double().cpu().numpy() for g in gs])}
    return res, check


def batched_bench(args, dev, stream) -> dict:
    """Batched uneven instances: meshes with triangle counts log-uniform in [200, 200,000]
    (tori and displaced spheres under random affine maps) and query counts log-uniform in
    [64, 1,048,576] uniform in each mesh's 1.5x box (drawn on the device from a seeded generator);
    Taylor order 1, scale 2.  Sequential single-mesh moment refits + query batches against the
    tuned batch replayed as one graph."""
    import torch

    from smallgwn_amd import gwn, synth

    rng = np.random.default_rng(2024)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2024)
    t0 = time.time()
    items, tris, nq = [], 0, 0
    geoms = []
    for k in range(args.batched_meshes):
        target = float(np.exp(rng.uniform(np.log(200), np.log(200_000))))
        if k % 2 == 0:
            nu = max(3, int(np.sqrt(target / 2)))
            base = synth.torus(nu, max(3, int(target / (2 * nu))))
        else:
            nlon = max(4, int(np.sqrt(target / 2)))
            base = synth.displaced_sphere(nlon, max(3, int(target / (2 * nlon))), seed=k)
        rot, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        v = (np.stack([base.vx, base.vy, base.vz], 1).astype(np.float64) @ rot.T * rng.uniform(0.5, 2.0)
             + rng.uniform(-5, 5, size=3)).astype(np.float32)
        geom = gwn.GeometryObject()
        gwn.gwn_upload_geometry(geom, v[:, 0], v[:, 1], v[:, 2], base.i0, base.i1, base.i2).raise_if_error()
        bvh = gwn.BvhObject()
        gwn.gwn_build_bvh(geom, bvh, gwn.BuildOptions(method=gwn.BuildMethod.HPLOC)).raise_if_error()
        mom = gwn.MomentObject(1)
        gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
        n = int(np.exp(rng.uniform(np.log(64), np.log(1 << 20))))
        lo, hi = v.min(0), v.max(0)
        c, e = (lo + hi) / 2, (hi - lo) / 2 * 1.5
        u = torch.rand((3, n), generator=gen, device=dev)
        q = [(torch.tensor(float(c[j] - e[j]), device=dev) + u[j] * float(2 * e[j])).contiguous() for j in range(3)]
        out = torch.empty(n, dtype=torch.float32, device=dev)
        items.append((bvh, mom, q[0], q[1], q[2], out))
        geoms.append(geom)
        tris += base.triangle_count
        nq += n
        if (k + 1) % 1024 == 0:
            log(f"batched setup: {k + 1} meshes ({time.time() - t0:.1f}s)")
    torch.cuda.synchronize()
    log(f"batched: {len(items)} meshes, {tris} tris, {nq} queries ({time.time() - t0:.1f}s setup)")

    def sequential():
        for bvh, mom, x, y, z, out in items:
            gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
            gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, x, y, z, out, accuracy_scale=2.0).raise_if_error()

    sequential()
    torch.cuda.synchronize()
    reps = 2
    t = time.perf_counter()
    for _ in range(reps):
        sequential()
    torch.cuda.synchronize()
    seq_ms = (time.perf_counter() - t) / reps * 1e3
    ref = [it[5].clone() for it in items]

    # Work of every mesh's queries (the batch takes the single-mesh decisions) for the roofline.
    c_tot = np.zeros(3, np.int64)
    for bvh, mom, x, y, z, out in items:
        cnt = torch.zeros(3 * x.numel(), dtype=torch.int32, device=dev)
        gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, x, y, z, out, accuracy_scale=2.0,
                                                    counters=cnt).raise_if_error()
        c_tot += cnt.view(-1, 3).to(torch.int64).sum(0).cpu().numpy()
    flops = float(c_tot[0] * FLOPS_CHILD_TEST + c_tot[1] * FLOPS_ACCEPT[1] + c_tot[2] * FLOPS_TRIANGLE)

    batch = gwn.BatchObject(items, accuracy_scale=2.0)
    tune_ms = batch.tune()
    batch.capture(include_refit=True).raise_if_error()
    batch.replay().raise_if_error()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        batch.replay().raise_if_error()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t) / reps * 1e3
    st = batch.status()
    identical = all(torch.equal(a, it[5]) for a, it in zip(ref, items))
    max_diff = max(float((a.double() - it[5].double()).abs().max()) if a.numel() else 0.0
                   for a, it in zip(ref, items))
    log(f"batched: sequential {seq_ms:.1f} ms, graph {graph_ms:.1f} ms ({seq_ms / graph_ms:.2f}x), "
        f"identical={identical}, max |diff| {max_diff:.2e}, tuning {batch.tuning()}")
    tuning = batch.tuning()
    batch.close()
    with_builds = batched_builds_leg(args, dev, geoms, items, reps)
    return {"meshes": len(items), "triangles": tris, "queries": nq, "taylor_order": 1, "builder": "hploc_r8",
            "sequential_ms": seq_ms, "graph_ms": graph_ms, "speedup": seq_ms / graph_ms,
            "mqueries_per_s": nq / (graph_ms * 1e-3) / 1e6, "best_query_launch_ms": tune_ms,
            "tuning": tuning, "bit_identical_to_sequential": bool(identical),
            "max_abs_diff_vs_sequential": max_diff, "refit_status_ok": st.is_ok(),
            "work": {"child_tests": int(c_tot[0]), "accepted": int(c_tot[1]), "near_field_triangles": int(c_tot[2])},
            "roofline": _valu_roofline(flops, tune_ms, "taylor_batch_kernel<1> (persistent, tuned)",
                                       "per-query work counters of every mesh (10 / 34 / 87 FLOP), "
                                       "over the tuned query launch"),
            "step": "moment refit of every mesh + Taylor queries of every mesh",
            "with_builds": with_builds}


def batched_builds_leg(args, dev, geoms, items, reps) -> dict:
    """The batched step with every tree rebuilt: a batch over the same geometries owns the trees
    and its graph replays all builds, moment refits and queries, against sequential gwn_build_bvh +
    gwn_refit_bvh_moment + Taylor batch calls with the same builder (H-PLOC r8, then LBVH; width 4)."""
    import torch

    from smallgwn_amd import gwn

    t0 = time.time()
    outs = [torch.empty_like(it[5]) for it in items]
    built = gwn.BuiltBatchObject([(g, it[2], it[3], it[4], o) for g, it, o in zip(geoms, items, outs)],
                                 width=4, hploc_radius=8, order=1, accuracy_scale=2.0)
    create_s = time.time() - t0
    query_ms = built.tune()
    result = {"create_s": create_s, "best_query_launch_ms": query_ms,
              "step": "build + moment refit of every mesh + Taylor queries of every mesh (one graph)"}
    if args.batched_tune_build:
        best = built.tune_build()
        cfg = built.build_config()
        result["tuned_build"] = {"best_ms": best, "method": cfg["method"].name.lower(), "width": cfg["width"],
                                 "hploc_search_radius": cfg["hploc_search_radius"]}
    for name, method in (("hploc_r8", gwn.BuildMethod.HPLOC), ("lbvh", gwn.BuildMethod.LBVH)):
        built.set_build(4, 8, method=method).raise_if_error()
        built.capture(include_refit=True).raise_if_error()
        built.replay().raise_if_error()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            built.replay().raise_if_error()
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - t) / reps * 1e3
        ok = built.status().is_ok()
        built.build().raise_if_error()
        torch.cuda.synchronize()
        t = time.perf_counter()
        built.build().raise_if_error()
        torch.cuda.synchronize()
        pass_ms = (time.perf_counter() - t) * 1e3
        opts = gwn.BuildOptions(method=method, hploc_search_radius=8, width=4)

        def sequential():
            for g, (bvh, mom, x, y, z, out) in zip(geoms, items):
                gwn.gwn_build_bvh(g, bvh, opts).raise_if_error()
                gwn.gwn_refit_bvh_moment(bvh, mom).raise_if_error()
                gwn.gwn_compute_winding_number_taylor_batch(bvh, mom, x, y, z, out,
                                                            accuracy_scale=2.0).raise_if_error()

        sequential()
        torch.cuda.synchronize()
        t = time.perf_counter()
        sequential()
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t) * 1e3
        identical = all(torch.equal(o, it[5]) for o, it in zip(outs, items))
        levels = built.build_config()["level_launches"]
        log(f"batched with {name} builds: sequential {seq_ms:.1f} ms, graph {graph_ms:.1f} ms "
            f"({seq_ms / graph_ms:.2f}x), builds + refits {pass_ms:.1f} ms, identical={identical}, "
            f"{levels} wide level launches (create {create_s:.1f}s)")
        result[name] = {"sequential_ms": seq_ms, "graph_ms": graph_ms, "speedup": seq_ms / graph_ms,
                        "builds_and_refits_ms": pass_ms, "level_launches": levels, "width": 4,
                        "bit_identical_to_sequential": bool(identical), "status_ok": ok}
    built.close()
    return result


def _profile_lines(name: str) -> list[str]:
    try:
        with open(os.path.join(REPO, "profiles", name)) as f:
            return f.readlines()
    except OSError:
        return []


def _profile_field(lines, key) -> str | None:
    for line in lines:
        if line.startswith(key + ":"):
            return line.split(":", 1)[1].strip()
    return None


def traversal_source_digest() -> str:
    """sha256 over the device source taylor_wave_kernel compiles from: the layout and winding
    device headers and gwn_winding.hip from the wave traversal's comment to leaf_root_kernel.  A
    counter summary records the digest it was measured at; a different digest means the kernel
    changed since, and its figures are left out of the line instead of being reported stale."""
    h = hashlib.sha256()
    for rel in ("include/gwn/detail/gwn_device_layout.hpp", "include/gwn/detail/gwn_device_winding.hpp"):
        with open(os.path.join(REPO, rel), encoding="utf-8") as f:
            h.update(f.read().encode())
    with open(os.path.join(REPO, "smallgwn_amd", "csrc", "gwn_winding.hip"), encoding="utf-8") as f:
        text = f.read()
    a = text.index("// Wave-cooperative traversal for spatially sorted queries.")
    b = text.index("template <int Prec, bool Grad>\n__global__ __launch_bounds__(256) void leaf_root_kernel(")
    h.update(text[a:b].encode())
    return h.hexdigest()


def _profile_source(name: str, lines) -> dict:
    """Where a counter summary's figures come from, and whether it still describes the kernel."""
    if not lines:
        return {}
    recorded = _profile_field(lines, "kernel_source_sha256")
    current = traversal_source_digest()
    return {"source": f"profiles/{name} (commit {_profile_field(lines, 'measured_at_commit')})",
            "current": recorded == current}


def valu_profile() -> dict:
    """VALU counters of the traversal kernel from the committed rocprofv3 --pmc summary (left out
    when the kernel's source digest differs from the one the summary was measured at)."""
    name = "taylor_wave_o2_large_valu.md"
    lines = _profile_lines(name)
    out = _profile_source(name, lines)
    if out.get("current"):
        for key in ("valu_lane_utilization", "valu_insts_per_wave"):
            v = _profile_field(lines, key)
            if v is not None:
                out[key] = float(v)
    return out


def traffic_profile() -> dict:
    """HBM bytes per traversal launch, from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE + WRITE_SIZE of the traversal kernel, scaled by the same counters on a
    256 MiB copy); left out when the kernel's source digest differs."""
    name = "taylor_wave_o2_large_traffic.md"
    lines = _profile_lines(name)
    out = _profile_source(name, lines)
    if out.get("current"):
        v = _profile_field(lines, "hbm_bytes_per_launch")
        if v is not None:
            out["hbm_bytes_per_launch"] = int(v)
    return out


def _cpu_threads() -> int:
    """Host threads for the CPU baseline: the cores this process may run on, at most 16 (the host
    share of one GPU on the MI355X boxes)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _median_runs(fn, runs=5):
    fn()  # warm-up
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), times


def cpu_baseline(mesh, qx, qy, qz, args, gpu_value) -> dict:
    """The oracle's C replay of the headline batch on the same H-PLOC radius-8 tree: Taylor order 2,
    the same accuracy scale, FP64 near-field triangles, over a strided query sample; one warm-up
    and five timed runs, median."""
    from oracle import oracle

    threads = _cpu_threads()
    t = time.perf_counter()
    bvh = oracle.build_bvh4(mesh, "hploc", 8)
    mom = oracle.refit_moment(bvh, args.order)
    build_s = time.perf_counter() - t
    stack = max(64, bvh.internal_stack_bound)

    def run(idx):
        oracle.winding_taylor_f32(bvh, mom, args.order, qx[idx], qy[idx], qz[idx], accuracy_scale=args.beta,
                                  stack_capacity=stack, threads=threads, near_fp64=True)

    probe_idx = np.linspace(0, qx.shape[0] - 1, 8192).astype(np.int64)
    t = time.perf_counter()
    run(probe_idx)
    probe = time.perf_counter() - t
    m = int(min(qx.shape[0], max(8192, 8192 * args.cpu_seconds / max(probe, 1e-3))))
    idx = np.linspace(0, qx.shape[0] - 1, m).astype(np.int64)
    med, times = _median_runs(lambda: run(idx))
    value = m / med / 1e6
    log(f"cpu baseline: {m} queries, median {med:.2f}s of {len(times)} runs on {threads} threads "
        f"({value:.4f} Mq/s; GPU/CPU {gpu_value / value:.0f}x)")
    return {
        "value": round(value, 5),
        "unit": "Mqueries/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "host_logical_cpus": os.cpu_count(),
        "runs_s": [round(x, 3) for x in times],
        "sample": f"{m} strided queries of the same 16,777,216-query workload on the same H-PLOC radius-8 "
                  f"tree (built by the oracle in {build_s:.1f}s, bit-identical to the GPU's), Taylor "
                  f"o{args.order}, beta={args.beta}, near-field triangles in FP64 (atan2 included); one "
                  f"warm-up and {len(times)} timed runs, median",
    }


def cpu_small() -> dict:
    """The small configuration fully on the CPU (the oracle): LBVH build and order-1 moments, exact
    winding (5,120 x 65,536 triangle evaluations), Taylor order 1 at scale 2, unsigned distance and
    65,536 first-hit rays; one warm-up and three timed runs each, median."""
    from oracle import oracle

    from smallgwn_amd import synth

    threads = _cpu_threads()
    mesh, q, rays = synth.config_small()
    bvh = oracle.build_bvh4(mesh)
    mom = oracle.refit_moment(bvh, 1)
    res = {"cores": threads, "kind": "port"}
    for name, fn in (
            ("build_and_refit", lambda: oracle.refit_moment(oracle.build_bvh4(mesh), 1)),
            ("exact", lambda: oracle.winding_exact_records_f32(bvh.triangles, *q, threads=threads)),
            ("taylor_o1", lambda: oracle.winding_taylor_f32(bvh, mom, 1, *q, accuracy_scale=2.0, threads=threads)),
            ("distance", lambda: oracle.unsigned_distance_f32(bvh, *q, threads=threads)),
            ("ray_first_hit", lambda: oracle.ray_first_hit_f32(bvh, *rays, threads=threads))):
        med, _ = _median_runs(fn, runs=3)
        res[name] = {"ms": med * 1e3} if name == "build_and_refit" else \
            {"ms": med * 1e3, "mqueries_per_s": q[0].shape[0] / med / 1e6}
    log("cpu small config: " + ", ".join(f"{k} {v['ms']:.1f} ms" for k, v in res.items() if isinstance(v, dict)))
    return res


def adjoint_gradient_check(check) -> dict:
    """dL/dV of the winding loss on the differentiation step's 4,096-query subsample: the GPU
    adjoint against the oracle's FP64 adjoint on the same tree (same decisions), relative error in
    the 2-norm and the largest entry error relative to the largest entry."""
    from oracle import oracle

    from smallgwn_amd import synth

    f = check["faces"]
    # The GPU tree was built on the initial positions and refit (bounds, then moments) after each
    # step: the oracle does the same, so both take the same accept/descend decisions.
    mesh = synth._mesh_from_arrays(check["vertices"], f)
    rb = oracle.refit_bvh(oracle.build_bvh4(synth._mesh_from_arrays(check["initial"], f), "hploc", 8), mesh)
    rm = oracle.refit_moment(rb, 1)
    rdv, _ = oracle.winding_taylor_adjoint_f64(rb, rm, 1, 2.0, mesh, *check["queries"], check["u"],
                                               threads=_cpu_threads())
    g = check["gpu"]
    rel = float(np.linalg.norm(g - rdv) / max(np.linalg.norm(rdv), 1e-300))
    ent = float(np.abs(g - rdv).max() / max(np.abs(rdv).max(), 1e-300))
    log(f"adjoint gradient vs FP64 oracle: relative error {rel:.2e} (max entry {ent:.2e})")
    return {"rel_error": rel, "max_entry_error_rel": ent, "queries": int(check["u"].shape[0]),
            "reference": "oracle FP64 adjoint on the same refit H-PLOC tree (pinned by central differences)"}


if __name__ == "__main__":
    sys.exit(main())
