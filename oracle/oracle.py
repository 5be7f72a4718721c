"""ctypes bindings for the CPU oracle (libgwn_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product path in
``smallgwn_amd`` never imports this module.
"""

from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgwn_oracle.so")

CHILD_DTYPE = np.dtype([("bounds", np.float32, (6,)), ("reference", np.uint64)])
NODE_DTYPE = np.dtype([("children", CHILD_DTYPE, (4,))])
assert CHILD_DTYPE.itemsize == 32 and NODE_DTYPE.itemsize == 128

OFFSET_MASK = (1 << 47) - 1
VALID_BIT = 1 << 47


class _Child(ctypes.Structure):
    _fields_ = [("bounds", ctypes.c_float * 6), ("reference", ctypes.c_uint64)]


class _Bvh(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_void_p),
        ("primitive_indices", ctypes.c_void_p),
        ("triangles", ctypes.c_void_p),
        ("sorted_morton", ctypes.c_void_p),
        ("binary_nodes", ctypes.c_void_p),
        ("binary_bounds", ctypes.c_void_p),
        ("root", _Child),
        ("node_count", ctypes.c_uint64),
        ("internal_stack_bound", ctypes.c_uint32),
        ("packed_stack_bound", ctypes.c_uint32),
        ("scene", ctypes.c_float * 9),
        ("binary_first", ctypes.c_void_p),
        ("binary_root", ctypes.c_uint32),
    ]


def build_library(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-s", "-C", _HERE]
        if force:
            cmd.append("-B")
        subprocess.run(cmd, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build_library()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
        L.gwn_oracle_morton_u32.restype = ctypes.c_uint32
        L.gwn_oracle_morton_u32.argtypes = [f32, f32, f32]
        L.gwn_oracle_morton_u64.restype = ctypes.c_uint64
        L.gwn_oracle_morton_u64.argtypes = [f32, f32, f32]
        L.gwn_oracle_stable_order_u64.argtypes = [vp, u64, vp]
        L.gwn_oracle_build_bvh4_f32.restype = i32
        L.gwn_oracle_build_bvh4_f32.argtypes = [vp, vp, vp, u64, vp, vp, vp, u64, ctypes.POINTER(_Bvh)]
        L.gwn_oracle_winding_taylor_frozen_f64.restype = ctypes.c_double
        L.gwn_oracle_winding_taylor_frozen_f64.argtypes = [
            vp, u64, _Child, vp, u64, vp, i32, f32, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, u64, i32]
        L.gwn_oracle_winding_taylor_adjoint_f64.restype = i32
        L.gwn_oracle_winding_taylor_adjoint_f64.argtypes = [
            vp, u64, _Child, vp, u64, vp, i32, f32, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp, vp, i32]
        L.gwn_oracle_distance_adjoint_f64.argtypes = [vp, u64, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp, vp, vp, i32]
        L.gwn_oracle_adjoint_magnitude_f64.argtypes = [vp, u64, vp, vp, vp, u64, vp, vp, vp, vp, u64, vp, vp, i32]
        L.gwn_oracle_build_bvh4_ex_f32.restype = i32
        L.gwn_oracle_build_bvh4_ex_f32.argtypes = [vp, vp, vp, u64, vp, vp, vp, u64, i32, ctypes.c_uint32,
                                                   ctypes.POINTER(_Bvh)]
        L.gwn_oracle_build_bvh_ex_f32.restype = i32
        L.gwn_oracle_build_bvh_ex_f32.argtypes = [vp, vp, vp, u64, vp, vp, vp, u64, i32, ctypes.c_uint32, i32,
                                                  ctypes.POINTER(_Bvh)]
        L.gwn_oracle_refit_bvh_f32.restype = i32
        L.gwn_oracle_refit_bvh_f32.argtypes = [vp, u64, ctypes.POINTER(_Child), vp, vp, u64, vp, vp, vp, u64, vp, vp, vp]
        L.gwn_oracle_moment_field_count.restype = i32
        L.gwn_oracle_moment_field_count.argtypes = [i32]
        L.gwn_oracle_refit_moment_f32.restype = i32
        L.gwn_oracle_refit_moment_f32.argtypes = [vp, u64, vp, u64, i32, vp]
        L.gwn_oracle_refit_moment_f64.restype = i32
        L.gwn_oracle_refit_moment_f64.argtypes = [vp, u64, vp, u64, i32, vp]
        L.gwn_oracle_winding_exact_records_f32.argtypes = [vp, u64, vp, vp, vp, u64, vp, i32]
        L.gwn_oracle_winding_exact_mesh_f64.argtypes = [vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, u64, vp, i32]
        L.gwn_oracle_solid_angle_f64.restype = ctypes.c_double
        L.gwn_oracle_solid_angle_f64.argtypes = [vp, vp, vp, vp]
        L.gwn_oracle_solid_angle_f32.restype = ctypes.c_float
        L.gwn_oracle_solid_angle_f32.argtypes = [vp, vp, vp, vp]
        L.gwn_oracle_triangle_payload_f64.argtypes = [vp, i32, vp]
        L.gwn_oracle_triangle_payload_f32.argtypes = [vp, i32, vp]
        L.gwn_oracle_winding_taylor_f32.argtypes = [
            vp, u64, _Child, vp, u64, vp, i32, f32, i32, vp, vp, vp, u64, vp, vp, i32, i32,
        ]
        L.gwn_oracle_winding_gradient_taylor_f32.argtypes = [
            vp, u64, _Child, vp, u64, vp, i32, f32, i32, vp, vp, vp, u64, vp, vp, vp, vp, vp, i32, i32,
        ]
        L.gwn_oracle_unsigned_distance_f32.argtypes = [vp, u64, _Child, vp, u64, vp, vp, vp, u64, f32, i32, vp, i32]
        L.gwn_oracle_ray_first_hit_f32.argtypes = [
            vp, u64, _Child, vp, vp, u64, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, u64, f32, f32, i32, i32, vp, i32,
        ]
        L.gwn_oracle_boundary_chain.restype = i32
        L.gwn_oracle_boundary_chain.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp]
        L.gwn_oracle_winding_antipodal_f32.argtypes = [
            vp, u64, _Child, vp, u64, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, u64, i32, i32, vp, vp, vp, vp, i32,
        ]
        L.gwn_oracle_point_triangle_distance2_f32.restype = f32
        L.gwn_oracle_point_triangle_distance2_f32.argtypes = [vp, vp, vp, vp]
        L.gwn_oracle_winding_gradient_exact_mesh_f64.argtypes = [
            vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, u64, vp, vp, vp, i32,
        ]
        L.gwn_oracle_winding_exact_mesh_d64.argtypes = [vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, u64, vp, i32]
        L.gwn_oracle_refit_moment_d64.restype = i32
        L.gwn_oracle_refit_moment_d64.argtypes = [vp, u64, vp, u64, vp, i32, vp]
        L.gwn_oracle_winding_taylor_d64.argtypes = [
            vp, u64, u64, vp, u64, vp, i32, ctypes.c_double, i32, vp, vp, vp, u64, vp, i32,
        ]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _c(a, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def morton_u64(nx: float, ny: float, nz: float) -> int:
    return int(lib().gwn_oracle_morton_u64(nx, ny, nz))


def morton_u32(nx: float, ny: float, nz: float) -> int:
    return int(lib().gwn_oracle_morton_u32(nx, ny, nz))


def stable_order_u64(keys) -> np.ndarray:
    k = _c(keys, np.uint64)
    out = np.empty(k.shape[0], dtype=np.uint32)
    lib().gwn_oracle_stable_order_u64(_ptr(k), k.shape[0], _ptr(out))
    return out


@dataclass
class OracleBvh:
    nodes: np.ndarray              # NODE_DTYPE[node_count]
    primitive_indices: np.ndarray  # uint32[T]
    triangles: np.ndarray          # float32[T, 9]  (v0, v1, v2)
    sorted_morton: np.ndarray      # uint64[T]
    binary_nodes: np.ndarray       # uint32[T-1, 4]  left, right, left_kind, right_kind
    binary_bounds: np.ndarray      # float32[T-1, 6]
    root_bounds: np.ndarray        # float32[6]
    root_reference: int
    internal_stack_bound: int
    packed_stack_bound: int
    scene: np.ndarray              # float32[9]
    binary_first: np.ndarray = None  # uint32[T-1]  smallest sorted leaf under each binary node
    binary_root: int = 0

    @property
    def node_count(self) -> int:
        return int(self.nodes.shape[0])


BUILD_HPLOC, BUILD_LBVH = 0, 1


def build_bvh4(mesh, method: str = "lbvh", radius: int = 8, width: int = 4) -> OracleBvh:
    """Wide BVH (width 2, 3 or 4, stored four-wide) of `mesh` with the LBVH or H-PLOC binary
    builder (H-PLOC search radius 1-8)."""
    vx, vy, vz = _c(mesh.vx, np.float32), _c(mesh.vy, np.float32), _c(mesh.vz, np.float32)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    T = i0.shape[0]
    nodes = np.zeros(max(T - 1, 1), dtype=NODE_DTYPE)
    prim = np.zeros(max(T, 1), dtype=np.uint32)
    tris = np.zeros((max(T, 1), 9), dtype=np.float32)
    morton = np.zeros(max(T, 1), dtype=np.uint64)
    bnodes = np.zeros((max(T - 1, 1), 4), dtype=np.uint32)
    bbounds = np.zeros((max(T - 1, 1), 6), dtype=np.float32)
    bfirst = np.zeros(max(T - 1, 1), dtype=np.uint32)
    out = _Bvh()
    out.nodes, out.primitive_indices, out.triangles = _ptr(nodes), _ptr(prim), _ptr(tris)
    out.sorted_morton, out.binary_nodes, out.binary_bounds = _ptr(morton), _ptr(bnodes), _ptr(bbounds)
    out.binary_first = _ptr(bfirst)
    m = {"lbvh": BUILD_LBVH, "hploc": BUILD_HPLOC}[method]
    st = lib().gwn_oracle_build_bvh_ex_f32(
        _ptr(vx), _ptr(vy), _ptr(vz), vx.shape[0], _ptr(i0), _ptr(i1), _ptr(i2), T, m, int(radius), int(width),
        ctypes.byref(out)
    )
    if st != 0:
        raise RuntimeError(f"oracle build failed with status {st}")
    n = int(out.node_count)
    return OracleBvh(
        nodes=nodes[:n].copy(),
        primitive_indices=prim[:T].copy(),
        triangles=tris[:T].copy(),
        sorted_morton=morton[:T].copy(),
        binary_nodes=bnodes[: max(T - 1, 0)].copy(),
        binary_bounds=bbounds[: max(T - 1, 0)].copy(),
        root_bounds=np.array(list(out.root.bounds), dtype=np.float32),
        root_reference=int(out.root.reference),
        internal_stack_bound=int(out.internal_stack_bound),
        packed_stack_bound=int(out.packed_stack_bound),
        scene=np.array(list(out.scene), dtype=np.float32),
        binary_first=bfirst[: max(T - 1, 0)].copy(),
        binary_root=int(out.binary_root),
    )


def refit_bvh(bvh: OracleBvh, mesh) -> OracleBvh:
    """Bounds refit of `bvh` against `mesh` (same triangle indexing): a new OracleBvh whose
    triangle records, child boxes and root box come from the mesh; topology is unchanged."""
    nodes = bvh.nodes.copy()
    tris = np.ascontiguousarray(bvh.triangles.copy())
    prim = _c(bvh.primitive_indices, np.uint32)
    vx, vy, vz = _c(mesh.vx, np.float32), _c(mesh.vy, np.float32), _c(mesh.vz, np.float32)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    st = lib().gwn_oracle_refit_bvh_f32(
        _ptr(nodes) if nodes.size else None, bvh.node_count, ctypes.byref(root), _ptr(prim), _ptr(tris),
        prim.shape[0], _ptr(vx), _ptr(vy), _ptr(vz), vx.shape[0], _ptr(i0), _ptr(i1), _ptr(i2),
    )
    if st != 0:
        raise RuntimeError(f"oracle bounds refit failed with status {st}")
    return dataclasses.replace(
        bvh, nodes=nodes, triangles=tris, root_bounds=np.array(list(root.bounds), dtype=np.float32)
    )


def field_count(order: int) -> int:
    return int(lib().gwn_oracle_moment_field_count(order))


def refit_moment(bvh: OracleBvh, order: int, dtype=np.float32) -> np.ndarray:
    """Moments in the reference's field-SoA node layout: array[node, field, slot]."""
    F = field_count(order)
    out = np.zeros((bvh.node_count, F, 4), dtype=dtype)
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    fn = lib().gwn_oracle_refit_moment_f32 if dtype == np.float32 else lib().gwn_oracle_refit_moment_f64
    st = fn(_ptr(nodes), bvh.node_count, _ptr(tris), tris.shape[0], order, _ptr(out))
    if st != 0:
        raise RuntimeError(f"oracle moment refit failed with status {st}")
    return out


def winding_exact_records_f32(triangles, qx, qy, qz, threads: int | None = None) -> np.ndarray:
    tris = _c(triangles, np.float32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    out = np.empty(qx.shape[0], dtype=np.float32)
    lib().gwn_oracle_winding_exact_records_f32(
        _ptr(tris), tris.shape[0], _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(out),
        threads or default_threads(),
    )
    return out


def winding_exact_mesh_f64(mesh, qx, qy, qz, threads: int | None = None) -> np.ndarray:
    vx, vy, vz = _c(mesh.vx, np.float32), _c(mesh.vy, np.float32), _c(mesh.vz, np.float32)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    out = np.empty(qx.shape[0], dtype=np.float64)
    lib().gwn_oracle_winding_exact_mesh_f64(
        _ptr(vx), _ptr(vy), _ptr(vz), _ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0],
        _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(out), threads or default_threads(),
    )
    return out


def winding_exact_mesh_d64(vertices, i0, i1, i2, qx, qy, qz, threads: int | None = None) -> np.ndarray:
    """Exact FP64 winding of a mesh with double vertices ``vertices`` (3, V): no float rounding
    anywhere (tests/reference_cpu.cuh:51-82 with Real = double)."""
    v = _c(np.asarray(vertices, dtype=np.float64).reshape(3, -1), np.float64)
    vx, vy, vz = _c(v[0], np.float64), _c(v[1], np.float64), _c(v[2], np.float64)
    i0, i1, i2 = _c(i0, np.uint32), _c(i1, np.uint32), _c(i2, np.uint32)
    qx, qy, qz = _c(qx, np.float64), _c(qy, np.float64), _c(qz, np.float64)
    out = np.empty(qx.shape[0], dtype=np.float64)
    lib().gwn_oracle_winding_exact_mesh_d64(
        _ptr(vx), _ptr(vy), _ptr(vz), _ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0],
        _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(out), threads or default_threads(),
    )
    return out


def double_tree(bvh: OracleBvh, vertices, i0, i1, i2):
    """Real = double records and boxes on the float-built tree ``bvh``: records (T, 9) of the double
    vertices ``vertices`` (3, V) in ``bvh``'s leaf order, every child's double box (N, 4, 6; min
    xyz, max xyz; zeros for empty slots), and the root box (6,).  Boxes are exact unions of the
    double triangles (gwn_bvh_refit_impl.cuh:23-352 with Real = double), folded in slot order."""
    v = np.asarray(vertices, dtype=np.float64).reshape(3, -1)
    idx = np.stack([np.asarray(i, dtype=np.int64) for i in (i0, i1, i2)])[:, bvh.primitive_indices]
    rec = np.concatenate([v[:, idx[0]].T, v[:, idx[1]].T, v[:, idx[2]].T], axis=1)
    corners = rec.reshape(-1, 3, 3)
    tri_box = np.concatenate([corners.min(axis=1), corners.max(axis=1)], axis=1)
    N = bvh.node_count
    child = np.zeros((N, 4, 6), dtype=np.float64)
    node_box = np.zeros((N, 6), dtype=np.float64)

    def leaf_box(ref):
        off, cnt = ref & ((1 << 47) - 1), ref >> 48
        b = tri_box[off:off + cnt]
        return np.concatenate([b[:, :3].min(axis=0), b[:, 3:].max(axis=0)])

    for n in range(N - 1, -1, -1):
        acc = None
        for s in range(4):
            ref = int(bvh.nodes[n]["children"][s]["reference"])
            if not ref >> 47 & 1:
                continue
            box = node_box[ref & ((1 << 47) - 1)] if ref >> 48 == 0 else leaf_box(ref)
            child[n, s] = box
            acc = box.copy() if acc is None else np.concatenate([np.minimum(acc[:3], box[:3]),
                                                                 np.maximum(acc[3:], box[3:])])
        node_box[n] = acc
    root = int(bvh.root_reference)
    root_box = node_box[0] if root >> 48 == 0 else leaf_box(root)
    return np.ascontiguousarray(rec), child, root_box


def refit_moment_d64(bvh: OracleBvh, triangles_d, child_bounds_d, order: int) -> np.ndarray:
    """Real = double moments [node, field, 4] from double records and double child boxes."""
    F = field_count(order)
    out = np.zeros((bvh.node_count, F, 4), dtype=np.float64)
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(triangles_d, np.float64)
    bounds = _c(child_bounds_d, np.float64)
    st = lib().gwn_oracle_refit_moment_d64(_ptr(nodes), bvh.node_count, _ptr(tris), tris.shape[0], _ptr(bounds),
                                           order, _ptr(out))
    if st != 0:
        raise RuntimeError(f"oracle double moment refit failed with status {st}")
    return out


def winding_taylor_d64(bvh: OracleBvh, triangles_d, moments_d, order: int, qx, qy, qz,
                       accuracy_scale: float = 2.0, stack_capacity: int = 64,
                       threads: int | None = None) -> np.ndarray:
    """Real = double Taylor traversal: the child references of ``bvh``, double records
    ``triangles_d`` (T, 9) and double moments ``moments_d`` [node, field, 4]."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(triangles_d, np.float64)
    mom = _c(moments_d, np.float64)
    qx, qy, qz = _c(qx, np.float64), _c(qy, np.float64), _c(qz, np.float64)
    out = np.empty(qx.shape[0], dtype=np.float64)
    lib().gwn_oracle_winding_taylor_d64(
        _ptr(nodes), bvh.node_count, bvh.root_reference, _ptr(tris), tris.shape[0], _ptr(mom), order,
        float(accuracy_scale), stack_capacity, _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(out),
        threads or default_threads(),
    )
    return out


def solid_angle_f64(a, b, c, q) -> float:
    a, b, c, q = (_c(v, np.float64) for v in (a, b, c, q))
    return float(lib().gwn_oracle_solid_angle_f64(_ptr(a), _ptr(b), _ptr(c), _ptr(q)))


def solid_angle_f32(a, b, c, q) -> float:
    a, b, c, q = (_c(v, np.float32) for v in (a, b, c, q))
    return float(lib().gwn_oracle_solid_angle_f32(_ptr(a), _ptr(b), _ptr(c), _ptr(q)))


def triangle_payload(record, order: int, dtype=np.float64) -> np.ndarray:
    """34 additive payload values of one (v0, v1, v2) record (see gwn_oracle.h)."""
    rec = _c(record, np.float32)
    out = np.zeros(34, dtype=dtype)
    fn = lib().gwn_oracle_triangle_payload_f64 if dtype == np.float64 else lib().gwn_oracle_triangle_payload_f32
    fn(_ptr(rec), order, _ptr(out))
    return out


def winding_taylor_f32(bvh: OracleBvh, moments: np.ndarray, order: int, qx, qy, qz,
                       accuracy_scale: float = 2.0, stack_capacity: int = 64,
                       counters: bool = False, threads: int | None = None, near_fp64: bool = False):
    """Replay of the Taylor traversal; ``near_fp64`` evaluates leaves in FP64 (precision split)."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    mom = _c(moments, np.float32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    out = np.empty(qx.shape[0], dtype=np.float32)
    cnt = np.zeros((qx.shape[0], 3), dtype=np.uint64) if counters else None
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    lib().gwn_oracle_winding_taylor_f32(
        _ptr(nodes), bvh.node_count, root, _ptr(tris), tris.shape[0], _ptr(mom), order,
        accuracy_scale, stack_capacity, _ptr(qx), _ptr(qy), _ptr(qz), qx.shape[0], _ptr(out),
        _ptr(cnt) if cnt is not None else None, 1 if near_fp64 else 0, threads or default_threads(),
    )
    return (out, cnt) if counters else out


def winding_gradient_taylor_f32(bvh: OracleBvh, moments: np.ndarray, order: int, qx, qy, qz,
                                accuracy_scale: float = 2.0, stack_capacity: int = 64,
                                counters: bool = False, threads: int | None = None,
                                near_fp64: bool = False):
    """Replay of the Taylor traversal with the analytic gradient: (winding, gx, gy, gz[, counters])."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    mom = _c(moments, np.float32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    n = qx.shape[0]
    out = np.empty(n, dtype=np.float32)
    g = [np.empty(n, dtype=np.float32) for _ in range(3)]
    cnt = np.zeros((n, 3), dtype=np.uint64) if counters else None
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    lib().gwn_oracle_winding_gradient_taylor_f32(
        _ptr(nodes), bvh.node_count, root, _ptr(tris), tris.shape[0], _ptr(mom), order,
        accuracy_scale, stack_capacity, _ptr(qx), _ptr(qy), _ptr(qz), n, _ptr(out),
        _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(cnt) if cnt is not None else None,
        1 if near_fp64 else 0, threads or default_threads(),
    )
    return (out, *g, cnt) if counters else (out, *g)


def winding_gradient_exact_mesh_f64(mesh, qx, qy, qz, threads: int | None = None):
    """Exact gradient of the winding number in FP64 over all triangles: (gx, gy, gz)."""
    vx, vy, vz = _c(mesh.vx, np.float32), _c(mesh.vy, np.float32), _c(mesh.vz, np.float32)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    n = qx.shape[0]
    g = [np.empty(n, dtype=np.float64) for _ in range(3)]
    lib().gwn_oracle_winding_gradient_exact_mesh_f64(
        _ptr(vx), _ptr(vy), _ptr(vz), _ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0],
        _ptr(qx), _ptr(qy), _ptr(qz), n, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), threads or default_threads(),
    )
    return tuple(g)


def unsigned_distance_f32(bvh: OracleBvh, qx, qy, qz, culling_band: float = float("inf"),
                          stack_capacity: int = 64, threads: int | None = None) -> np.ndarray:
    """Replay of the near-first unsigned-distance traversal in FP32."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    out = np.empty(qx.shape[0], dtype=np.float32)
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    lib().gwn_oracle_unsigned_distance_f32(
        _ptr(nodes), bvh.node_count, root, _ptr(tris), tris.shape[0], _ptr(qx), _ptr(qy), _ptr(qz),
        qx.shape[0], culling_band, stack_capacity, _ptr(out), threads or default_threads(),
    )
    return out


# gwn_ray_first_hit_result<float, uint32_t> (include/gwn/gwn_query.cuh:74-95), 32 bytes.
RAY_HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("nx", np.float32),
                          ("ny", np.float32), ("nz", np.float32), ("primitive_id", np.uint32),
                          ("status", np.uint8), ("pad", np.uint8, (3,))])
assert RAY_HIT_DTYPE.itemsize == 32


def ray_first_hit_f32(bvh: OracleBvh, ox, oy, oz, dx, dy, dz, t_min: float = 0.0, t_max: float = float("inf"),
                      stack_capacity: int = 64, threads: int | None = None, traversal: str = "auto") -> np.ndarray:
    """Replay of the batch ray first-hit query: a RAY_HIT_DTYPE record per ray.  ``traversal``:
    "auto" (the batch's dispatch), "canonical" or "packed"."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    prim = _c(bvh.primitive_indices, np.uint32)
    arrs = [_c(a, np.float32) for a in (ox, oy, oz, dx, dy, dz)]
    out = np.zeros(arrs[0].shape[0], dtype=RAY_HIT_DTYPE)
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    lib().gwn_oracle_ray_first_hit_f32(
        _ptr(nodes), bvh.node_count, root, _ptr(tris), _ptr(prim), tris.shape[0], bvh.packed_stack_bound,
        *[_ptr(a) for a in arrs], arrs[0].shape[0], t_min, t_max, stack_capacity,
        {"auto": -1, "canonical": 0, "packed": 1}[traversal], _ptr(out), threads or default_threads(),
    )
    return out


def boundary_chain(mesh):
    """Directed boundary rows of the oriented triangle chain: (start, end, multiplicity)."""
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    m = max(3 * i0.shape[0], 1)
    start = np.zeros(m, np.uint32)
    end = np.zeros(m, np.uint32)
    mult = np.zeros(m, np.uint64)
    rows = ctypes.c_uint64(0)
    st = lib().gwn_oracle_boundary_chain(_ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0], _ptr(start), _ptr(end),
                                         _ptr(mult), ctypes.byref(rows))
    if st != 0:
        raise RuntimeError(f"oracle boundary chain failed with status {st}")
    r = rows.value
    return start[:r].copy(), end[:r].copy(), mult[:r].copy()


def winding_antipodal_f32(bvh: OracleBvh, mesh, boundary, qx, qy, qz, gradient: bool = False,
                          near_fp64: bool = False, stack_capacity: int = 64, threads: int | None = None):
    """Antipodal winding number (and, with ``gradient``, (w, gx, gy, gz)) of the oriented mesh;
    ``boundary`` is the (start, end, multiplicity) triple of ``boundary_chain``."""
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    vx, vy, vz = _c(mesh.vx, np.float32), _c(mesh.vy, np.float32), _c(mesh.vz, np.float32)
    bs, be, bm = _c(boundary[0], np.uint32), _c(boundary[1], np.uint32), _c(boundary[2], np.uint64)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    n = qx.shape[0]
    out = np.empty(n, np.float32)
    g = [np.empty(n, np.float32) for _ in range(3)] if gradient else [None, None, None]
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    lib().gwn_oracle_winding_antipodal_f32(
        _ptr(nodes), bvh.node_count, root, _ptr(tris), tris.shape[0], _ptr(vx), _ptr(vy), _ptr(vz),
        _ptr(bs) if bs.size else None, _ptr(be) if be.size else None, _ptr(bm) if bm.size else None, bs.shape[0],
        _ptr(qx), _ptr(qy), _ptr(qz), n, stack_capacity, 1 if near_fp64 else 0, _ptr(out),
        *[(_ptr(a) if a is not None else None) for a in g], threads or default_threads(),
    )
    return (out, *g) if gradient else out


def point_triangle_distance2_f32(p, a, b, c) -> float:
    p, a, b, c = (_c(np.asarray(v), np.float32) for v in (p, a, b, c))
    return float(lib().gwn_oracle_point_triangle_distance2_f32(_ptr(p), _ptr(a), _ptr(b), _ptr(c)))


def canonical_dfs(nodes: np.ndarray, root_reference: int):
    """Slot-order DFS serialization: (kind, offset-or-leaf, count, bounds) with internal offsets erased."""
    out = []
    if not (root_reference & VALID_BIT):
        return out
    if (root_reference >> 48) != 0:
        return [("leaf", root_reference & OFFSET_MASK, root_reference >> 48)]
    stack = [root_reference & OFFSET_MASK]
    while stack:
        n = stack.pop()
        kids = []
        for s in range(4):
            ch = nodes[n]["children"][s]
            ref = int(ch["reference"])
            b = tuple(np.asarray(ch["bounds"]).view(np.uint32).tolist())
            if not ref & VALID_BIT:
                out.append(("invalid", s))
                continue
            if (ref >> 48) != 0:
                out.append(("leaf", s, ref & OFFSET_MASK, ref >> 48, b))
            else:
                out.append(("internal", s, b))
                kids.append(ref & OFFSET_MASK)
        stack.extend(reversed(kids))
    return out


# ---------------------------------------------------------------------------
# Adjoint (no reference counterpart): FP64 derivatives of L = sum u_q w_q and L = sum u_p d_p.
# ---------------------------------------------------------------------------

def _vertices64(mesh, vertices):
    if vertices is None:
        vertices = np.stack([mesh.vx, mesh.vy, mesh.vz]).astype(np.float64)
    return _c(np.asarray(vertices, dtype=np.float64).reshape(3, -1), np.float64)


def _adj_args(bvh, moments, order, beta, mesh, vertices, qx, qy, qz, u):
    nodes = _c(bvh.nodes, NODE_DTYPE)
    tris = _c(bvh.triangles, np.float32)
    mom = _c(moments, np.float32)
    prim = _c(bvh.primitive_indices, np.uint32)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    v = _vertices64(mesh, vertices)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    u = _c(u, np.float64)
    root = _Child()
    for k in range(6):
        root.bounds[k] = float(bvh.root_bounds[k])
    root.reference = bvh.root_reference
    keep = (nodes, tris, mom, prim, i0, i1, i2, v, qx, qy, qz, u)
    args = [_ptr(nodes), bvh.node_count, root, _ptr(tris), tris.shape[0], _ptr(mom), int(order), float(beta),
            _ptr(prim), _ptr(i0), _ptr(i1), _ptr(i2), _ptr(v), v.shape[1], _ptr(qx), _ptr(qy), _ptr(qz),
            _ptr(u), qx.shape[0]]
    return keep, args


def winding_taylor_frozen_f64(bvh, moments, order, beta, mesh, qx, qy, qz, u, vertices=None, threads=None) -> float:
    """sum_q u_q w_q in FP64 for vertex positions `vertices` (3, V) with the accept/descend decisions
    the FP32 forward takes on `bvh`/`moments` (the function whose derivative the adjoint gives)."""
    keep, args = _adj_args(bvh, moments, order, beta, mesh, vertices, qx, qy, qz, u)
    return float(lib().gwn_oracle_winding_taylor_frozen_f64(*args, threads or default_threads()))


def winding_taylor_adjoint_f64(bvh, moments, order, beta, mesh, qx, qy, qz, u, vertices=None, threads=None):
    """(dL/dV as float64 (3, V), dL/dq as float64 (Q, 3)) for L = sum_q u_q w_q, Taylor orders 0-1."""
    keep, args = _adj_args(bvh, moments, order, beta, mesh, vertices, qx, qy, qz, u)
    V = keep[7].shape[1]
    dv = np.zeros((3, V), np.float64)
    dq = np.zeros((keep[8].shape[0], 3), np.float64)
    st = lib().gwn_oracle_winding_taylor_adjoint_f64(*args, _ptr(dv), _ptr(dq), threads or default_threads())
    if st != 0:
        raise RuntimeError(f"oracle adjoint failed with status {st}")
    return dv, dq


def adjoint_magnitude_f64(mesh, qx, qy, qz, u, vertices=None, threads=None):
    """Error envelope of the winding adjoint: (mag_dv (3, V), mag_dq (Q, 3)), the sums over all
    triangles and queries of |u_q| |d(W_t(q) / 4 pi)/dv| at each vertex and of |u_q| |dW_t/dq| per
    query, every term exact in FP64 (brute force, no BVH)."""
    v = _vertices64(mesh, vertices)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    qx, qy, qz = _c(qx, np.float32), _c(qy, np.float32), _c(qz, np.float32)
    u = _c(u, np.float64)
    n = qx.shape[0]
    mv = np.zeros((3, v.shape[1]), np.float64)
    mq = np.zeros((n, 3), np.float64)
    lib().gwn_oracle_adjoint_magnitude_f64(_ptr(v), v.shape[1], _ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0],
                                           _ptr(qx), _ptr(qy), _ptr(qz), _ptr(u), n, _ptr(mv), _ptr(mq),
                                           threads or default_threads())
    return mv, mq


def distance_adjoint_f64(mesh, px, py, pz, u, vertices=None, threads=None):
    """Brute-force FP64 distances and (dist, dL/dV (3, V), dL/dp (P, 3)) for L = sum_p u_p d_p."""
    v = _vertices64(mesh, vertices)
    i0, i1, i2 = _c(mesh.i0, np.uint32), _c(mesh.i1, np.uint32), _c(mesh.i2, np.uint32)
    px, py, pz = _c(px, np.float32), _c(py, np.float32), _c(pz, np.float32)
    u = _c(u, np.float64)
    n = px.shape[0]
    dist = np.zeros(n, np.float64)
    dv = np.zeros((3, v.shape[1]), np.float64)
    dp = np.zeros((n, 3), np.float64)
    lib().gwn_oracle_distance_adjoint_f64(_ptr(v), v.shape[1], _ptr(i0), _ptr(i1), _ptr(i2), i0.shape[0],
                                          _ptr(px), _ptr(py), _ptr(pz), _ptr(u), n, _ptr(dist), _ptr(dv),
                                          _ptr(dp), threads or default_threads())
    return dist, dv, dp
