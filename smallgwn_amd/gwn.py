"""Python mirror of the reference's gwn_* API over libgwn_hip.so (ctypes, C ABI).

Names, argument meaning and error behaviour follow the reference's header-only C++ API
(kririae/smallgwn include/gwn/*.cuh): objects own device storage, functions return a
``Status`` instead of raising (``gwn_status``), batch queries take device arrays and
enqueue on a stream without synchronizing.  Device arrays are torch tensors on a
ROCm device (any object exposing ``data_ptr()`` works); streams are torch streams or raw
``hipStream_t`` integers.

The HIP library must be built (``__graft_entry__.build()``); importing this module
without it raises ImportError -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgwn_hip.so")

DEFAULT_BLOCK_SIZE = 256          # include/gwn/gwn_query.cuh:42
DEFAULT_STACK_CAPACITY = 64       # include/gwn/gwn_query.cuh:45

EXPORTED_SYMBOLS = (
    "gwn_abi_version", "gwn_last_error_message",
    "gwn_geometry_create", "gwn_geometry_destroy", "gwn_upload_geometry_f32",
    "gwn_update_geometry_f32", "gwn_update_geometry_device_f32", "gwn_geometry_counts",
    "gwn_bvh_create", "gwn_bvh_destroy", "gwn_build_bvh4_f32", "gwn_build_bvh_f32", "gwn_refit_bvh_f32",
    "gwn_bvh_get_info",
    "gwn_bvh_download",
    "gwn_boundary_create", "gwn_boundary_destroy", "gwn_build_boundary_chain", "gwn_boundary_info",
    "gwn_boundary_download",
    "gwn_moment_create", "gwn_moment_destroy", "gwn_refit_bvh_moment_f32", "gwn_moment_download",
    "gwn_moment_order",
    "gwn_compute_winding_number_exact_batch_f32", "gwn_compute_winding_number_taylor_batch_f32",
    "gwn_compute_winding_number_exact_batch_ex_f32", "gwn_compute_winding_number_taylor_batch_ex_f32",
    "gwn_compute_winding_number_taylor_batch_noalloc_f32",
    "gwn_compute_winding_gradient_taylor_batch_f32", "gwn_compute_winding_gradient_taylor_batch_ex_f32",
    "gwn_compute_unsigned_distance_batch_f32", "gwn_compute_ray_first_hit_batch_f32",
    "gwn_compute_winding_number_antipodal_batch_f32", "gwn_compute_winding_number_antipodal_batch_ex_f32",
    "gwn_compute_winding_gradient_antipodal_batch_f32", "gwn_compute_winding_gradient_antipodal_batch_ex_f32",
    "gwn_compute_winding_taylor_adjoint_batch_f32", "gwn_compute_unsigned_distance_adjoint_batch_f32",
    "gwn_compute_unsigned_distance_batch_ex_f32", "gwn_compute_unsigned_distance_adjoint_batch_ex_f32",
    "gwn_batch_create", "gwn_batch_destroy", "gwn_batch_refit_moments", "gwn_batch_query",
    "gwn_batch_set_tuning", "gwn_batch_get_tuning", "gwn_batch_tune", "gwn_batch_capture", "gwn_batch_replay",
    "gwn_batch_status", "gwn_batch_create_built", "gwn_batch_build", "gwn_batch_set_build", "gwn_batch_get_build",
    "gwn_batch_tune_build", "gwn_batch_export", "gwn_compute_unsigned_distance_work_f32", "gwn_compute_ray_first_hit_work_f32",
    "gwn_compute_query_order_f32", "gwn_radix_sort_pairs_u64", "gwn_radix_sort_pairs_u32",
    "gwn_upload_geometry_f32_u64", "gwn_upload_geometry_f64", "gwn_upload_geometry_f64_u64",
    "gwn_update_geometry_f64", "gwn_update_geometry_device_f64", "gwn_build_bvh_f64", "gwn_refit_bvh_f64",
    "gwn_bvh_download_f64", "gwn_refit_bvh_moment_f64", "gwn_moment_download_f64",
    "gwn_compute_winding_number_exact_batch_f64", "gwn_compute_winding_number_taylor_batch_f64",
    "gwn_moment_info",
    *(f"gwn_{obj}_{op}" for obj in ("geometry", "bvh", "moment", "boundary")
      for op in ("stream", "set_stream", "clear", "get_accessor")),
)


class Error(enum.IntEnum):
    """gwn_error (include/gwn/gwn_utils.cuh:213-218); cuda_runtime_error maps to the HIP runtime."""

    SUCCESS = 0
    INVALID_ARGUMENT = 1
    HIP_RUNTIME_ERROR = 2
    INTERNAL_ERROR = 3


class BuildMethod(enum.IntEnum):
    """gwn_bvh_build_method (include/gwn/gwn_bvh_build.cuh:16-20)."""

    HPLOC = 0
    LBVH = 1


class Precision(enum.IntEnum):
    PLAIN = 0   # the reference's float arithmetic
    SPLIT = 1   # FP32 traversal / far field, FP64 near-field triangle terms
    FP64 = 2    # every term in FP64 (exact batch only)


@dataclass(frozen=True)
class Status:
    """gwn_status: an error category and a diagnostic message."""

    error: Error = Error.SUCCESS
    message: str = ""

    def is_ok(self) -> bool:
        return self.error == Error.SUCCESS

    def raise_if_error(self) -> "Status":
        if not self.is_ok():
            raise GwnError(self)
        return self


class GwnError(RuntimeError):
    def __init__(self, status: Status):
        super().__init__(f"{status.error.name}: {status.message}")
        self.status = status


@dataclass(frozen=True)
class BuildOptions:
    """gwn_bvh_build_options (include/gwn/gwn_bvh_build.cuh:22-27)."""

    method: BuildMethod = BuildMethod.HPLOC
    hploc_search_radius: int = 8
    width: int = 4  # the Width template argument of gwn_bvh_object<Width, ...>: 2, 3 or 4


@dataclass(frozen=True)
class QueryConfig:
    """gwn_query_batch_config (include/gwn/gwn_query.cuh:48-52) plus the query presort switch."""

    block_size: int = DEFAULT_BLOCK_SIZE
    stack_capacity: int = DEFAULT_STACK_CAPACITY
    sort_queries: int = -1
    lanes_per_point: int = 0  # unsigned distance: 0 (four) or 1 or 4 lanes per point, same results


class _CConfig(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("stack_capacity", ctypes.c_int32),
                ("sort_queries", ctypes.c_int32), ("lanes_per_point", ctypes.c_int32)]


class _CBvhInfo(ctypes.Structure):
    _fields_ = [
        ("node_count", ctypes.c_uint64), ("triangle_count", ctypes.c_uint64), ("revision", ctypes.c_uint64),
        ("internal_stack_bound", ctypes.c_uint32), ("packed_stack_bound", ctypes.c_uint32),
        ("max_depth", ctypes.c_uint32), ("width", ctypes.c_uint32),
        ("root_bounds", ctypes.c_float * 6), ("root_reference", ctypes.c_uint64),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)"
        )
    # torch first, so the library binds to the HIP runtime torch already loaded.
    import torch  # noqa: F401

    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i32, u32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float
    f64 = ctypes.c_double
    sigs = {
        "gwn_abi_version": [],
        "gwn_geometry_create": [vp],
        "gwn_geometry_destroy": [vp],
        "gwn_upload_geometry_f32": [vp, vp, vp, vp, u64, vp, vp, vp, u64, vp],
        "gwn_update_geometry_f32": [vp, vp, vp, vp, u64, vp],
        "gwn_update_geometry_device_f32": [vp, vp, vp, vp, u64, vp],
        "gwn_geometry_counts": [vp, vp, vp],
        "gwn_bvh_create": [vp],
        "gwn_bvh_destroy": [vp],
        "gwn_build_bvh4_f32": [vp, vp, i32, u32, vp],
        "gwn_build_bvh_f32": [vp, vp, i32, i32, u32, vp],
        "gwn_refit_bvh_f32": [vp, vp, vp],
        "gwn_bvh_get_info": [vp, vp],
        "gwn_bvh_download": [vp, vp, vp, vp, vp],
        "gwn_boundary_create": [vp],
        "gwn_boundary_destroy": [vp],
        "gwn_build_boundary_chain": [vp, vp, vp],
        "gwn_boundary_info": [vp, vp, vp, vp, vp],
        "gwn_boundary_download": [vp, vp, vp, vp, vp],
        "gwn_moment_create": [vp, i32],
        "gwn_moment_destroy": [vp],
        "gwn_refit_bvh_moment_f32": [vp, vp, vp],
        "gwn_moment_download": [vp, vp, vp],
        "gwn_moment_order": [vp, vp],
        "gwn_compute_winding_number_exact_batch_f32": [vp, vp, vp, vp, vp, u64, vp, vp],
        "gwn_compute_winding_number_taylor_batch_f32": [vp, vp, vp, vp, vp, vp, u64, f32, vp, vp],
        "gwn_compute_winding_number_exact_batch_ex_f32": [vp, vp, vp, vp, vp, u64, i32, vp, vp],
        "gwn_compute_winding_number_taylor_batch_ex_f32": [vp, vp, vp, vp, vp, vp, u64, f32, i32, vp, vp, vp, vp],
        "gwn_compute_winding_number_taylor_batch_noalloc_f32": [vp, vp, vp, vp, vp, vp, u64, f32, i32, vp, vp],
        "gwn_compute_winding_gradient_taylor_batch_f32": [vp, vp, vp, vp, vp, vp, vp, vp, u64, f32, vp, vp],
        "gwn_compute_winding_gradient_taylor_batch_ex_f32": [
            vp, vp, vp, vp, vp, vp, vp, vp, vp, u64, f32, i32, vp, vp, vp, vp],
        "gwn_compute_unsigned_distance_batch_f32": [vp, vp, vp, vp, vp, u64, f32, vp, vp],
        "gwn_compute_ray_first_hit_batch_f32": [vp, vp, vp, vp, vp, vp, vp, vp, u64, f32, f32, vp, vp],
        "gwn_compute_winding_number_antipodal_batch_f32": [vp, vp, vp, vp, vp, vp, vp, u64, vp, vp],
        "gwn_compute_winding_number_antipodal_batch_ex_f32": [vp, vp, vp, vp, vp, vp, vp, u64, i32, vp, vp],
        "gwn_compute_winding_gradient_antipodal_batch_f32": [vp, vp, vp, vp, vp, vp, vp, vp, u64, vp, vp],
        "gwn_compute_winding_gradient_antipodal_batch_ex_f32": [vp, vp, vp, vp, vp, vp, vp, vp, u64, i32, vp, vp],
        "gwn_compute_winding_taylor_adjoint_batch_f32": [
            vp, vp, vp, vp, vp, vp, vp, u64, f32, vp, vp, vp, vp, vp, vp, vp, vp],
        "gwn_compute_unsigned_distance_adjoint_batch_f32": [
            vp, vp, vp, vp, vp, vp, u64, f32, vp, vp, vp, vp, vp, vp, vp, vp],
        "gwn_compute_unsigned_distance_batch_ex_f32": [vp, vp, vp, vp, vp, vp, u64, f32, vp, vp],
        "gwn_compute_unsigned_distance_adjoint_batch_ex_f32": [
            vp, vp, vp, vp, vp, vp, vp, u64, vp, vp, vp, vp, vp, vp, vp],
        "gwn_batch_create": [vp, u64, f32, vp, vp],
        "gwn_batch_destroy": [vp],
        "gwn_batch_refit_moments": [vp, vp],
        "gwn_batch_query": [vp, vp],
        "gwn_batch_set_tuning": [vp, vp, vp],
        "gwn_batch_get_tuning": [vp, vp],
        "gwn_batch_tune": [vp, vp, vp],
        "gwn_batch_capture": [vp, i32],
        "gwn_batch_replay": [vp, vp],
        "gwn_batch_status": [vp, vp],
        "gwn_batch_create_built": [vp, u64, vp, f32, vp, vp],
        "gwn_batch_build": [vp, vp],
        "gwn_batch_set_build": [vp, vp, vp],
        "gwn_batch_get_build": [vp, vp, vp],
        "gwn_batch_tune_build": [vp, vp, vp],
        "gwn_batch_export": [vp, u64, vp, vp, vp],
        "gwn_compute_unsigned_distance_work_f32": [vp, vp, vp, vp, u64, f32, vp, vp],
        "gwn_compute_ray_first_hit_work_f32": [vp, vp, vp, vp, vp, vp, vp, u64, f32, f32, vp, vp, vp],
        "gwn_compute_query_order_f32": [vp, vp, vp, vp, vp, u64, vp],
        "gwn_radix_sort_pairs_u64": [vp, vp, u64, i32, i32, vp],
        "gwn_radix_sort_pairs_u32": [vp, vp, u64, i32, i32, vp],
        "gwn_upload_geometry_f32_u64": [vp, vp, vp, vp, u64, vp, vp, vp, u64, vp],
        "gwn_upload_geometry_f64": [vp, vp, vp, vp, u64, vp, vp, vp, u64, vp],
        "gwn_upload_geometry_f64_u64": [vp, vp, vp, vp, u64, vp, vp, vp, u64, vp],
        "gwn_update_geometry_f64": [vp, vp, vp, vp, u64, vp],
        "gwn_update_geometry_device_f64": [vp, vp, vp, vp, u64, vp],
        "gwn_build_bvh_f64": [vp, vp, i32, i32, u32, vp],
        "gwn_refit_bvh_f64": [vp, vp, vp],
        "gwn_bvh_download_f64": [vp, vp, vp, vp, vp],
        "gwn_refit_bvh_moment_f64": [vp, vp, vp],
        "gwn_moment_download_f64": [vp, vp, vp],
        "gwn_compute_winding_number_exact_batch_f64": [vp, vp, vp, vp, vp, u64, vp, vp],
        "gwn_compute_winding_number_taylor_batch_f64": [vp, vp, vp, vp, vp, vp, u64, f64, vp, vp],
        "gwn_moment_info": [vp, vp, vp],
    }
    for obj in ("geometry", "bvh", "moment", "boundary"):
        sigs[f"gwn_{obj}_stream"] = [vp, vp]
        sigs[f"gwn_{obj}_set_stream"] = [vp, vp]
        sigs[f"gwn_{obj}_clear"] = [vp]
        sigs[f"gwn_{obj}_get_accessor"] = [vp, vp]
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.gwn_last_error_message.argtypes = []
    L.gwn_last_error_message.restype = ctypes.c_char_p
    return L


_lib = _load()


def library() -> ctypes.CDLL:
    return _lib


def _status(code: int) -> Status:
    if code == 0:
        return Status()
    msg = _lib.gwn_last_error_message()
    return Status(Error(code), msg.decode() if msg else "")


def _stream(stream) -> int | None:
    if stream is None:
        import torch

        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _dptr(t) -> int | None:
    if t is None:
        return None
    return int(t.data_ptr())


def _config(cfg: QueryConfig | None) -> _CConfig:
    cfg = cfg or QueryConfig()
    return _CConfig(cfg.block_size, cfg.stack_capacity, cfg.sort_queries, cfg.lanes_per_point)


class _CGeometryAccessor(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("i0", ctypes.c_void_p), ("i1", ctypes.c_void_p), ("i2", ctypes.c_void_p),
                ("vertex_count", ctypes.c_uint64), ("triangle_count", ctypes.c_uint64)]


class _CBvhAccessor(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("node_count", ctypes.c_uint64),
                ("primitive_indices", ctypes.c_void_p), ("triangles", ctypes.c_void_p),
                ("triangle_count", ctypes.c_uint64), ("root_bounds", ctypes.c_float * 6),
                ("root_reference", ctypes.c_uint64), ("internal_stack_bound", ctypes.c_uint32),
                ("packed_stack_bound", ctypes.c_uint32), ("revision", ctypes.c_uint64)]


class _CMomentAccessor(ctypes.Structure):
    _fields_ = [("records", ctypes.c_void_p), ("node_count", ctypes.c_uint64), ("order", ctypes.c_int32),
                ("record_bytes", ctypes.c_uint32), ("width", ctypes.c_int32), ("bvh_revision", ctypes.c_uint64),
                ("eval_records", ctypes.c_void_p)]


class _CBoundaryAccessor(ctypes.Structure):
    _fields_ = [("start", ctypes.c_void_p), ("end", ctypes.c_void_p), ("multiplicity", ctypes.c_void_p),
                ("edge_count", ctypes.c_uint64), ("mesh_vertex_count", ctypes.c_uint64),
                ("mesh_triangle_count", ctypes.c_uint64), ("is_built", ctypes.c_int32)]


def _accessor_dict(acc) -> dict:
    out = {}
    for name, _ in acc._fields_:
        v = getattr(acc, name)
        if isinstance(v, ctypes.Array):
            v = np.array(list(v), dtype=np.float32)
        out[name] = v
    return out


class _Handle:
    _create = None
    _destroy = None

    def __init__(self, *args):
        h = ctypes.c_void_p()
        _status(getattr(_lib, self._create)(ctypes.byref(h), *args)).raise_if_error()
        self._h = h
        # The Real of the storage the last successful upload, build or refit produced (4 or 8
        # bytes); it selects the _f32 or _f64 entry points.
        self.real_bytes = 4

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    # Stream binding and release (include/gwn/gwn_utils.cuh:59-80): storage is released on the
    # bound stream; set_stream rebinds without synchronizing; clear releases and keeps the binding.
    def stream(self) -> int | None:
        """The bound stream's raw handle (None for the null stream)."""
        out = ctypes.c_void_p()
        _status(getattr(_lib, f"gwn_{self._kind}_stream")(self._h, ctypes.byref(out))).raise_if_error()
        return out.value

    def set_stream(self, stream) -> None:
        _status(getattr(_lib, f"gwn_{self._kind}_set_stream")(self._h, _stream(stream))).raise_if_error()

    def clear(self) -> None:
        _status(getattr(_lib, f"gwn_{self._kind}_clear")(self._h)).raise_if_error()

    def accessor(self) -> dict:
        """Non-owning snapshot of the device storage (raw device addresses as ints)."""
        acc = self._accessor_type()
        _status(getattr(_lib, f"gwn_{self._kind}_get_accessor")(self._h, ctypes.byref(acc))).raise_if_error()
        return _accessor_dict(acc)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            getattr(_lib, self._destroy)(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GeometryObject(_Handle):
    """gwn_geometry_object<float, uint32_t> (include/gwn/gwn_geometry.cuh:97-149)."""

    _kind = "geometry"
    _accessor_type = _CGeometryAccessor
    _create = "gwn_geometry_create"
    _destroy = "gwn_geometry_destroy"

    def counts(self) -> tuple[int, int]:
        v, t = ctypes.c_uint64(), ctypes.c_uint64()
        _status(_lib.gwn_geometry_counts(self._h, ctypes.byref(v), ctypes.byref(t))).raise_if_error()
        return int(v.value), int(t.value)

    def vertex_count(self) -> int:
        return self.counts()[0]

    def triangle_count(self) -> int:
        return self.counts()[1]

    def has_data(self) -> bool:
        v, t = self.counts()
        return v != 0 or t != 0


class BvhObject(_Handle):
    """gwn_bvh_object<4, float, uint32_t> (include/gwn/gwn_bvh.cuh:331-380)."""

    _kind = "bvh"
    _accessor_type = _CBvhAccessor
    _create = "gwn_bvh_create"
    _destroy = "gwn_bvh_destroy"

    def info(self) -> dict:
        info = _CBvhInfo()
        _status(_lib.gwn_bvh_get_info(self._h, ctypes.byref(info))).raise_if_error()
        return {
            "node_count": int(info.node_count),
            "triangle_count": int(info.triangle_count),
            "revision": int(info.revision),
            "internal_stack_bound": int(info.internal_stack_bound),
            "packed_stack_bound": int(info.packed_stack_bound),
            "max_depth": int(info.max_depth),
            "width": int(info.width),
            "root_bounds": np.array(list(info.root_bounds), dtype=np.float32),
            "root_reference": int(info.root_reference),
        }

    def has_data(self) -> bool:
        return self.info()["revision"] != 0

    def download(self, stream=None):
        """Host copies: nodes (reference gwn_bvh_node<Width, Real> bytes: 32 B per child for float,
        56 B for double), primitive order, records (t, 9) in the BVH's Real."""
        inf = self.info()
        n, t = inf["node_count"], inf["triangle_count"]
        dbl = self.real_bytes == 8
        nodes = np.zeros(n * (56 if dbl else 32) * inf["width"], dtype=np.uint8)
        prim = np.zeros(t, dtype=np.uint32)
        tris = np.zeros((t, 9), dtype=np.float64 if dbl else np.float32)
        _status((_lib.gwn_bvh_download_f64 if dbl else _lib.gwn_bvh_download)(
            self._h, nodes.ctypes.data if n else None, prim.ctypes.data if t else None,
            tris.ctypes.data if t else None, _stream(stream),
        )).raise_if_error()
        return nodes, prim, tris


class MomentObject(_Handle):
    """gwn_bvh_moment_object<4, Order, float, uint32_t> (include/gwn/gwn_bvh.cuh:383-433)."""

    _kind = "moment"
    _accessor_type = _CMomentAccessor
    _create = "gwn_moment_create"
    _destroy = "gwn_moment_destroy"

    def __init__(self, order: int):
        super().__init__(int(order))
        self.order = int(order)

    def download(self, node_count: int, stream=None) -> np.ndarray:
        """Coefficients in the reference field-SoA layout gwn_bvh_moment_node<Width, Order>:
        array[node, field, slot] for the width of the BVH they were formed from."""
        F = {0: 7, 1: 13, 2: 23}[self.order]
        acc = self.accessor()
        if node_count != acc["node_count"]:
            raise ValueError(f"moment holds {acc['node_count']} nodes, not {node_count}")
        out = np.zeros((node_count, F, acc["width"]), dtype=np.float32)
        if node_count:
            _status(_lib.gwn_moment_download(self._h, out.ctypes.data, _stream(stream))).raise_if_error()
        return out

    def download_f64(self, node_count: int, width: int, stream=None) -> np.ndarray:
        """Real = double coefficients, array[node, field, slot] (Width = the BVH's width)."""
        F = {0: 7, 1: 13, 2: 23}[self.order]
        out = np.zeros((node_count, F, width), dtype=np.float64)
        if node_count:
            _status(_lib.gwn_moment_download_f64(self._h, out.ctypes.data, _stream(stream))).raise_if_error()
        return out


class BoundaryChainObject(_Handle):
    """gwn_boundary_chain_object<uint32_t> (include/gwn/gwn_boundary.cuh:62-115)."""

    _kind = "boundary"
    _accessor_type = _CBoundaryAccessor
    _create = "gwn_boundary_create"
    _destroy = "gwn_boundary_destroy"

    def info(self) -> dict:
        e, v, t = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        b = ctypes.c_int32()
        _status(_lib.gwn_boundary_info(self._h, ctypes.byref(e), ctypes.byref(v), ctypes.byref(t),
                                       ctypes.byref(b))).raise_if_error()
        return {"edge_count": e.value, "mesh_vertex_count": v.value, "mesh_triangle_count": t.value,
                "is_built": bool(b.value)}

    def has_data(self) -> bool:
        return self.info()["is_built"]

    def download(self, stream=None):
        """Rows (start, end, multiplicity) as host arrays."""
        n = self.info()["edge_count"]
        start = np.zeros(n, np.uint32)
        end = np.zeros(n, np.uint32)
        mult = np.zeros(n, np.uint64)
        if n:
            _status(_lib.gwn_boundary_download(self._h, start.ctypes.data, end.ctypes.data, mult.ctypes.data,
                                               _stream(stream))).raise_if_error()
        return start, end, mult


# ---------------------------------------------------------------------------------------------
# free functions with the reference's names
# ---------------------------------------------------------------------------------------------

def _host(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def gwn_upload_geometry(geometry: GeometryObject, x, y, z, i0, i1, i2, stream=None, real=np.float32,
                        index=np.uint32) -> Status:
    """gwn_upload_geometry (include/gwn/gwn_geometry.cuh:66-71): host SoA arrays.

    ``real`` (np.float32 or np.float64) and ``index`` (np.uint32 or np.uint64) are the geometry
    object's Real and Index template arguments; the arrays are converted to them.
    """
    real, index = np.dtype(real), np.dtype(index)
    if real not in (np.float32, np.float64) or index not in (np.uint32, np.uint64):
        return Status(Error.INVALID_ARGUMENT, "Real must be float32/float64 and Index uint32/uint64.")
    x, y, z = _host(x, real), _host(y, real), _host(z, real)
    i0, i1, i2 = _host(i0, index), _host(i1, index), _host(i2, index)
    if not (x.shape == y.shape == z.shape):
        return Status(Error.INVALID_ARGUMENT, "Vertex SoA spans must have identical lengths.")
    if not (i0.shape == i1.shape == i2.shape):
        return Status(Error.INVALID_ARGUMENT, "Triangle SoA spans must have identical lengths.")
    nv, nt = x.shape[0], i0.shape[0]
    fn = {(4, 4): _lib.gwn_upload_geometry_f32, (4, 8): _lib.gwn_upload_geometry_f32_u64,
          (8, 4): _lib.gwn_upload_geometry_f64, (8, 8): _lib.gwn_upload_geometry_f64_u64}[
        (real.itemsize, index.itemsize)]
    st = _status(fn(
        geometry.handle, x.ctypes.data if nv else None, y.ctypes.data if nv else None,
        z.ctypes.data if nv else None, nv, i0.ctypes.data if nt else None,
        i1.ctypes.data if nt else None, i2.ctypes.data if nt else None, nt, _stream(stream),
    ))
    if st.is_ok():
        geometry.real_bytes = real.itemsize
    return st


def gwn_update_geometry(geometry: GeometryObject, x, y, z, stream=None) -> Status:
    """gwn_update_geometry (include/gwn/gwn_geometry.cuh:84-90); device tensors are accepted too."""
    dbl = geometry.real_bytes == 8
    if hasattr(x, "data_ptr") and getattr(x, "is_cuda", False):
        if not (x.numel() == y.numel() == z.numel()):
            return Status(Error.INVALID_ARGUMENT, "Vertex SoA spans must have identical lengths.")
        fn = _lib.gwn_update_geometry_device_f64 if dbl else _lib.gwn_update_geometry_device_f32
        return _status(fn(geometry.handle, _dptr(x), _dptr(y), _dptr(z), x.numel(), _stream(stream)))
    real = np.float64 if dbl else np.float32
    x, y, z = _host(x, real), _host(y, real), _host(z, real)
    if not (x.shape == y.shape == z.shape):
        return Status(Error.INVALID_ARGUMENT, "Vertex SoA spans must have identical lengths.")
    n = x.shape[0]
    return _status((_lib.gwn_update_geometry_f64 if dbl else _lib.gwn_update_geometry_f32)(
        geometry.handle, x.ctypes.data if n else None, y.ctypes.data if n else None,
        z.ctypes.data if n else None, n, _stream(stream)))


def gwn_build_bvh(geometry: GeometryObject, bvh: BvhObject, options: BuildOptions = BuildOptions(),
                  stream=None) -> Status:
    """gwn_build_bvh<Width, Real, Index> (include/gwn/gwn_bvh_build.cuh:34-39), Width = options.width,
    Real = the geometry's."""
    if geometry.real_bytes == 8:
        st = _status(_lib.gwn_build_bvh_f64(
            geometry.handle, bvh.handle, int(options.width), int(options.method),
            int(options.hploc_search_radius), _stream(stream)))
    elif options.width == 4:
        st = _status(_lib.gwn_build_bvh4_f32(
            geometry.handle, bvh.handle, int(options.method), int(options.hploc_search_radius), _stream(stream)))
    else:
        st = _status(_lib.gwn_build_bvh_f32(
            geometry.handle, bvh.handle, int(options.width), int(options.method),
            int(options.hploc_search_radius), _stream(stream)))
    if st.is_ok():
        bvh.real_bytes = geometry.real_bytes
    return st


def gwn_build_boundary_chain(geometry: GeometryObject, boundary: BoundaryChainObject, stream=None) -> Status:
    """gwn_build_boundary_chain (include/gwn/gwn_boundary.cuh:117-125)."""
    return _status(_lib.gwn_build_boundary_chain(geometry.handle, boundary.handle, _stream(stream)))


def gwn_refit_bvh(geometry: GeometryObject, bvh: BvhObject, stream=None) -> Status:
    """gwn_refit_bvh (include/gwn/gwn_bvh_refit.cuh:25-29): records and boxes from the updated
    geometry, same topology; the BVH gets a new revision, so moments need a refit."""
    fn = _lib.gwn_refit_bvh_f64 if bvh.real_bytes == 8 else _lib.gwn_refit_bvh_f32
    return _status(fn(geometry.handle, bvh.handle, _stream(stream)))


def gwn_refit_bvh_moment(bvh: BvhObject, moment: MomentObject, stream=None) -> Status:
    """gwn_refit_bvh_moment<Order> (include/gwn/gwn_bvh_refit.cuh:36-51)."""
    fn = _lib.gwn_refit_bvh_moment_f64 if bvh.real_bytes == 8 else _lib.gwn_refit_bvh_moment_f32
    st = _status(fn(bvh.handle, moment.handle, _stream(stream)))
    if st.is_ok():
        moment.real_bytes = bvh.real_bytes
    return st


def _numel(t) -> int:
    return 0 if t is None else int(t.numel())


def _check_soa(qx, qy, qz, outs) -> Status | None:
    n = _numel(qx)
    if _numel(qy) != n or _numel(qz) != n:
        return Status(Error.INVALID_ARGUMENT, "Query SoA spans must have identical lengths.")
    for o in outs:
        if _numel(o) != n:
            return Status(Error.INVALID_ARGUMENT, "Output span size must match query count.")
    return None


def _check_real(real_bytes: int, *tensors) -> Status | None:
    """Device spans must hold the object's Real (the kernels read them with that width)."""
    for t in tensors:
        if t is not None and getattr(t, "element_size", None) is not None and t.numel() \
                and (t.element_size() != real_bytes or not t.is_floating_point()):
            return Status(Error.INVALID_ARGUMENT,
                          f"Query/output spans must hold the objects' Real ({8 * real_bytes}-bit floats).")
    return None


def gwn_compute_winding_number_exact_batch(bvh: BvhObject, qx, qy, qz, out,
                                           config: QueryConfig | None = None, stream=None) -> Status:
    """gwn_compute_winding_number_exact_batch (include/gwn/gwn_query.cuh:308-346); Real = the BVH's."""
    bad = _check_soa(qx, qy, qz, [out]) or _check_real(bvh.real_bytes, qx, qy, qz, out)
    if bad:
        return bad
    cfg = _config(config)
    fn = (_lib.gwn_compute_winding_number_exact_batch_f64 if bvh.real_bytes == 8
          else _lib.gwn_compute_winding_number_exact_batch_f32)
    return _status(fn(
        bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_winding_number_exact_batch_ex(bvh: BvhObject, qx, qy, qz, out_f64, precision: Precision,
                                              config: QueryConfig | None = None, stream=None) -> Status:
    """Exact batch with a precision mode; ``out_f64`` is a float64 device array."""
    bad = _check_soa(qx, qy, qz, [out_f64])
    if bad:
        return bad
    cfg = _config(config)
    return _status(_lib.gwn_compute_winding_number_exact_batch_ex_f32(
        bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out_f64), _numel(qx), int(precision),
        ctypes.byref(cfg), _stream(stream)))


def gwn_compute_winding_number_taylor_batch(bvh: BvhObject, moment: MomentObject, qx, qy, qz, out,
                                            accuracy_scale: float = 2.0, config: QueryConfig | None = None,
                                            stream=None, precision: Precision = Precision.PLAIN,
                                            counters=None, order=None) -> Status:
    """gwn_compute_winding_number_taylor_batch<Order> (include/gwn/gwn_query.cuh:384-445).

    ``precision=Precision.SPLIT`` evaluates near-field triangles in FP64.  ``order`` (optional
    uint32 device array from ``gwn_compute_query_order``) fixes the visiting order.  ``counters``
    (optional uint32 device array of 3 * count) receives per-query child tests, accepted
    expansions and near-field triangle evaluations.
    """
    bad = _check_soa(qx, qy, qz, [out]) or _check_real(bvh.real_bytes, qx, qy, qz, out)
    if bad:
        return bad
    if order is not None and _numel(order) != _numel(qx):
        return Status(Error.INVALID_ARGUMENT, "Query order span size must match query count.")
    cfg = _config(config)
    if bvh.real_bytes == 8:
        if precision != Precision.PLAIN or counters is not None or order is not None:
            return Status(Error.INVALID_ARGUMENT,
                          "Real = double Taylor batches take no precision mode, counters or order.")
        return _status(_lib.gwn_compute_winding_number_taylor_batch_f64(
            bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx),
            float(accuracy_scale), ctypes.byref(cfg), _stream(stream)))
    if precision == Precision.PLAIN and counters is None and order is None:
        return _status(_lib.gwn_compute_winding_number_taylor_batch_f32(
            bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx),
            float(accuracy_scale), ctypes.byref(cfg), _stream(stream)))
    return _status(_lib.gwn_compute_winding_number_taylor_batch_ex_f32(
        bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx),
        float(accuracy_scale), int(precision), ctypes.byref(cfg), _dptr(order), _dptr(counters),
        _stream(stream)))


NOALLOC_MAX_QUERIES = 1 << 20


def gwn_compute_winding_number_taylor_batch_noalloc(bvh: BvhObject, moment: MomentObject, qx, qy, qz, out,
                                                    accuracy_scale: float = 2.0,
                                                    config: QueryConfig | None = None, stream=None,
                                                    precision: Precision = Precision.PLAIN) -> Status:
    """The Taylor winding batch without device allocation (no reference counterpart).

    Same results as ``gwn_compute_winding_number_taylor_batch``; the Hilbert presort's scratch is
    the library's static per-GPU array, so at most ``NOALLOC_MAX_QUERIES`` queries per call, and
    calls that overlap on one GPU race.
    """
    bad = _check_soa(qx, qy, qz, [out]) or _check_real(4, qx, qy, qz, out)
    if bad:
        return bad
    cfg = _config(config)
    return _status(_lib.gwn_compute_winding_number_taylor_batch_noalloc_f32(
        bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx),
        float(accuracy_scale), int(precision), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_winding_gradient_taylor_batch(bvh: BvhObject, moment: MomentObject, qx, qy, qz, gx, gy, gz,
                                              accuracy_scale: float = 2.0, config: QueryConfig | None = None,
                                              stream=None, precision: Precision = Precision.PLAIN,
                                              out=None, counters=None, order=None) -> Status:
    """gwn_compute_winding_gradient_taylor_batch<Order> (include/gwn/gwn_query.cuh:684-756).

    Writes the analytic gradient of the Taylor winding number to gx, gy, gz; ``out`` (optional)
    receives the winding number of the same traversal.  ``precision``, ``order`` and ``counters``
    as for ``gwn_compute_winding_number_taylor_batch``.
    """
    bad = _check_soa(qx, qy, qz, [gx, gy, gz] + ([out] if out is not None else []))
    if bad:
        return bad
    if order is not None and _numel(order) != _numel(qx):
        return Status(Error.INVALID_ARGUMENT, "Query order span size must match query count.")
    cfg = _config(config)
    if precision == Precision.PLAIN and out is None and counters is None and order is None:
        return _status(_lib.gwn_compute_winding_gradient_taylor_batch_f32(
            bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(gx), _dptr(gy), _dptr(gz),
            _numel(qx), float(accuracy_scale), ctypes.byref(cfg), _stream(stream)))
    return _status(_lib.gwn_compute_winding_gradient_taylor_batch_ex_f32(
        bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _dptr(gx), _dptr(gy), _dptr(gz),
        _numel(qx), float(accuracy_scale), int(precision), ctypes.byref(cfg), _dptr(order), _dptr(counters),
        _stream(stream)))


def gwn_compute_unsigned_distance_batch(bvh: BvhObject, qx, qy, qz, out, culling_band: float = float("inf"),
                                        config: QueryConfig | None = None, stream=None, nearest=None) -> Status:
    """gwn_compute_unsigned_distance_batch (include/gwn/gwn_query.cuh:129-180).

    ``nearest`` (optional uint32/int32 device array, one entry per query) receives each query's
    nearest triangle record for ``gwn_compute_unsigned_distance_adjoint_batch(..., nearest=...)``:
    0xFFFFFFFF where the distance is the band or zero, 0xFFFFFFFE on stack overflow."""
    bad = _check_soa(qx, qy, qz, [out] + ([nearest] if nearest is not None else []))
    if bad:
        return bad
    cfg = _config(config)
    if nearest is not None:
        return _status(_lib.gwn_compute_unsigned_distance_batch_ex_f32(
            bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _dptr(nearest), _numel(qx),
            float(culling_band), ctypes.byref(cfg), _stream(stream)))
    return _status(_lib.gwn_compute_unsigned_distance_batch_f32(
        bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), _numel(qx), float(culling_band),
        ctypes.byref(cfg), _stream(stream)))


# gwn_ray_first_hit_result<float, uint32_t> as a numpy dtype (32 bytes per record).
RAY_HIT_DTYPE = np.dtype([("t", np.float32), ("u", np.float32), ("v", np.float32), ("nx", np.float32),
                          ("ny", np.float32), ("nz", np.float32), ("primitive_id", np.uint32),
                          ("status", np.uint8), ("pad", np.uint8, (3,))])


class RayStatus(enum.IntEnum):
    """gwn_ray_first_hit_status (include/gwn/gwn_query.cuh:63-68)."""

    MISS = 0
    HIT = 1
    OVERFLOW = 2


def gwn_compute_ray_first_hit_batch(bvh: BvhObject, ox, oy, oz, dx, dy, dz, out, t_min: float = 0.0,
                                    t_max: float = float("inf"), config: QueryConfig | None = None,
                                    stream=None) -> Status:
    """gwn_compute_ray_first_hit_batch (include/gwn/gwn_query.cuh:205-290).

    ``out`` is a device buffer of 32 bytes per ray (e.g. a uint8 tensor of 32 * count or an int32
    tensor of 8 * count); view it on the host with ``RAY_HIT_DTYPE``."""
    n = _numel(ox)
    if any(_numel(a) != n for a in (oy, oz, dx, dy, dz)) or int(out.numel() * out.element_size()) != 32 * n:
        return Status(Error.INVALID_ARGUMENT, "ray first-hit: mismatched span sizes")
    cfg = _config(config)
    return _status(_lib.gwn_compute_ray_first_hit_batch_f32(
        bvh.handle, _dptr(ox), _dptr(oy), _dptr(oz), _dptr(dx), _dptr(dy), _dptr(dz), _dptr(out), n,
        float(t_min), float(t_max), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_unsigned_distance_work(bvh: BvhObject, qx, qy, qz, counters, culling_band: float = float("inf"),
                                       stream=None) -> Status:
    """Work counters of the unsigned-distance traversal: counters[2q] child box tests, counters[2q+1]
    point-triangle tests of query q (int32/uint32 device array of 2 x count; no reference counterpart)."""
    n = _numel(qx)
    if _numel(qy) != n or _numel(qz) != n or _numel(counters) != 2 * n:
        return Status(Error.INVALID_ARGUMENT, "Query spans and counters (2 per query) must match.")
    return _status(_lib.gwn_compute_unsigned_distance_work_f32(bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), n,
                                                               float(culling_band), _dptr(counters), _stream(stream)))


def gwn_compute_ray_first_hit_work(bvh: BvhObject, ox, oy, oz, dx, dy, dz, counters, t_min: float = 0.0,
                                   t_max: float = float("inf"), config: QueryConfig | None = None,
                                   stream=None) -> Status:
    """Work counters of the ray traversal: counters[2r] slab tests, counters[2r+1] triangle tests."""
    n = _numel(ox)
    if any(_numel(a) != n for a in (oy, oz, dx, dy, dz)) or _numel(counters) != 2 * n:
        return Status(Error.INVALID_ARGUMENT, "Ray spans and counters (2 per ray) must match.")
    c = _config(config)
    return _status(_lib.gwn_compute_ray_first_hit_work_f32(
        bvh.handle, _dptr(ox), _dptr(oy), _dptr(oz), _dptr(dx), _dptr(dy), _dptr(dz), n, float(t_min), float(t_max),
        ctypes.byref(c), _dptr(counters), _stream(stream)))


def gwn_compute_winding_number_antipodal_batch(geometry: GeometryObject, bvh: BvhObject,
                                               boundary: BoundaryChainObject, qx, qy, qz, out,
                                               config: QueryConfig | None = None, stream=None,
                                               precision: Precision = Precision.PLAIN) -> Status:
    """gwn_compute_winding_number_antipodal_batch (include/gwn/gwn_query.cuh:492-552).

    Positive-axis crossings over the BVH's triangles plus the boundary chain's spherical areas;
    axes retried +Z, +X, +Y, NaN when all three are singular.  ``precision=Precision.SPLIT``
    evaluates the boundary areas in FP64."""
    n = _numel(qx)
    if _numel(qy) != n or _numel(qz) != n:
        return Status(Error.INVALID_ARGUMENT, "Query SoA spans require identical lengths.")
    if _numel(out) != n:
        return Status(Error.INVALID_ARGUMENT, "Output span size differs from query count.")
    cfg = _config(config)
    if precision == Precision.PLAIN:
        return _status(_lib.gwn_compute_winding_number_antipodal_batch_f32(
            geometry.handle, bvh.handle, boundary.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), n,
            ctypes.byref(cfg), _stream(stream)))
    return _status(_lib.gwn_compute_winding_number_antipodal_batch_ex_f32(
        geometry.handle, bvh.handle, boundary.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), n,
        int(precision), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_winding_gradient_antipodal_batch(geometry: GeometryObject, boundary: BoundaryChainObject,
                                                 qx, qy, qz, gx, gy, gz, config: QueryConfig | None = None,
                                                 stream=None, precision: Precision = Precision.PLAIN) -> Status:
    """gwn_compute_winding_gradient_antipodal_batch (include/gwn/gwn_query.cuh:582-640)."""
    n = _numel(qx)
    if _numel(qy) != n or _numel(qz) != n:
        return Status(Error.INVALID_ARGUMENT, "Query SoA spans require identical lengths.")
    if _numel(gx) != n or _numel(gy) != n or _numel(gz) != n:
        return Status(Error.INVALID_ARGUMENT, "Output span sizes differ from query count.")
    cfg = _config(config)
    if precision == Precision.PLAIN:
        return _status(_lib.gwn_compute_winding_gradient_antipodal_batch_f32(
            geometry.handle, boundary.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(gx), _dptr(gy), _dptr(gz),
            n, ctypes.byref(cfg), _stream(stream)))
    return _status(_lib.gwn_compute_winding_gradient_antipodal_batch_ex_f32(
        geometry.handle, boundary.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(gx), _dptr(gy), _dptr(gz),
        n, int(precision), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_winding_taylor_adjoint_batch(geometry: GeometryObject, bvh: BvhObject, moment: MomentObject,
                                             qx, qy, qz, grad_w, grad_vx, grad_vy, grad_vz, grad_qx=None,
                                             grad_qy=None, grad_qz=None, accuracy_scale: float = 2.0,
                                             config: QueryConfig | None = None, stream=None) -> Status:
    """dL/dV (added to grad_v*) and dL/dq (written to grad_q* when given) of L = sum grad_w * w for the
    Taylor winding of orders 0 and 1 (no reference counterpart)."""
    n = _numel(qx)
    if _numel(qy) != n or _numel(qz) != n or _numel(grad_w) != n:
        return Status(Error.INVALID_ARGUMENT, "Query SoA spans require identical lengths.")
    if grad_qx is not None and any(_numel(a) != n for a in (grad_qx, grad_qy, grad_qz)):
        return Status(Error.INVALID_ARGUMENT, "Query gradient span sizes differ from query count.")
    nv = geometry.vertex_count()
    if any(_numel(a) != nv for a in (grad_vx, grad_vy, grad_vz)):
        return Status(Error.INVALID_ARGUMENT, "Vertex gradient span sizes differ from vertex count.")
    cfg = _config(config)
    return _status(_lib.gwn_compute_winding_taylor_adjoint_batch_f32(
        geometry.handle, bvh.handle, moment.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(grad_w), n,
        float(accuracy_scale), _dptr(grad_vx), _dptr(grad_vy), _dptr(grad_vz), _dptr(grad_qx), _dptr(grad_qy),
        _dptr(grad_qz), ctypes.byref(cfg), _stream(stream)))


def gwn_compute_unsigned_distance_adjoint_batch(geometry: GeometryObject, bvh: BvhObject, px, py, pz, grad_d,
                                                grad_vx, grad_vy, grad_vz, grad_px=None, grad_py=None,
                                                grad_pz=None, culling_band: float = float("inf"),
                                                config: QueryConfig | None = None, stream=None,
                                                nearest=None) -> Status:
    """dL/dV (added) and dL/dp (written when given) of L = sum grad_d * d for the unsigned distance.

    With ``nearest`` (the forward's records from ``gwn_compute_unsigned_distance_batch``) no
    traversal runs; ``culling_band`` and ``config`` are then unused."""
    n = _numel(px)
    if _numel(py) != n or _numel(pz) != n or _numel(grad_d) != n:
        return Status(Error.INVALID_ARGUMENT, "Query SoA spans require identical lengths.")
    if grad_px is not None and any(_numel(a) != n for a in (grad_px, grad_py, grad_pz)):
        return Status(Error.INVALID_ARGUMENT, "Point gradient span sizes differ from point count.")
    nv = geometry.vertex_count()
    if any(_numel(a) != nv for a in (grad_vx, grad_vy, grad_vz)):
        return Status(Error.INVALID_ARGUMENT, "Vertex gradient span sizes differ from vertex count.")
    if nearest is not None:
        if _numel(nearest) != n:
            return Status(Error.INVALID_ARGUMENT, "Nearest-record span size differs from point count.")
        return _status(_lib.gwn_compute_unsigned_distance_adjoint_batch_ex_f32(
            geometry.handle, bvh.handle, _dptr(px), _dptr(py), _dptr(pz), _dptr(nearest), _dptr(grad_d), n,
            _dptr(grad_vx), _dptr(grad_vy), _dptr(grad_vz), _dptr(grad_px), _dptr(grad_py), _dptr(grad_pz),
            _stream(stream)))
    cfg = _config(config)
    return _status(_lib.gwn_compute_unsigned_distance_adjoint_batch_f32(
        geometry.handle, bvh.handle, _dptr(px), _dptr(py), _dptr(pz), _dptr(grad_d), n, float(culling_band),
        _dptr(grad_vx), _dptr(grad_vy), _dptr(grad_vz), _dptr(grad_px), _dptr(grad_py), _dptr(grad_pz),
        ctypes.byref(cfg), _stream(stream)))


class _BatchItem(ctypes.Structure):
    _fields_ = [("bvh", ctypes.c_void_p), ("moment", ctypes.c_void_p), ("qx", ctypes.c_void_p),
                ("qy", ctypes.c_void_p), ("qz", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("count", ctypes.c_uint64)]


class _BatchTuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int32), ("chunk", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("traversal", ctypes.c_int32), ("stack_capacity", ctypes.c_int32),
                ("wave_queries_per_triangle", ctypes.c_float), ("meshes", ctypes.c_uint64),
                ("queries", ctypes.c_uint64)]


class BatchObject:
    """Batched mode (no reference counterpart): Taylor moment refits and Taylor winding queries of
    many meshes in persistent launches, replayable as one hipGraph.  ``items`` are tuples
    (bvh, moment, qx, qy, qz, out) of library objects and float32 device tensors; every moment has
    the same order.  The objects and tensors are kept alive by this object."""

    def __init__(self, items, accuracy_scale: float = 2.0, stream=None):
        self._keep = list(items)
        arr = (_BatchItem * max(1, len(self._keep)))()
        for k, (bvh, mom, qx, qy, qz, out) in enumerate(self._keep):
            n = _numel(qx)
            if _numel(qy) != n or _numel(qz) != n or _numel(out) != n:
                raise ValueError("Query SoA spans require identical lengths.")
            arr[k] = _BatchItem(bvh.handle.value, mom.handle.value, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), n)
        h = ctypes.c_void_p()
        _status(_lib.gwn_batch_create(arr, len(self._keep), float(accuracy_scale), ctypes.byref(h),
                                      _stream(stream))).raise_if_error()
        self._h = h

    @property
    def handle(self):
        return self._h

    def refit_moments(self, stream=None) -> Status:
        return _status(_lib.gwn_batch_refit_moments(self._h, _stream(stream)))

    def query(self, stream=None) -> Status:
        return _status(_lib.gwn_batch_query(self._h, _stream(stream)))

    def tune(self, stream=None) -> float:
        ms = ctypes.c_float()
        _status(_lib.gwn_batch_tune(self._h, _stream(stream), ctypes.byref(ms))).raise_if_error()
        return float(ms.value)

    def tuning(self) -> dict:
        t = _BatchTuning()
        _status(_lib.gwn_batch_get_tuning(self._h, ctypes.byref(t))).raise_if_error()
        return {k: (float if k == "wave_queries_per_triangle" else int)(getattr(t, k))
                for k, _ in _BatchTuning._fields_}

    def set_tuning(self, block_size: int, chunk: int, blocks_per_cu: int, traversal: int = 0,
                   stack_capacity: int = 0, wave_queries_per_triangle: float = 1.0, stream=None) -> Status:
        """traversal 0: one query per lane; 1: wave-cooperative over each mesh's Hilbert order;
        2: wave-cooperative for meshes with at least wave_queries_per_triangle queries per
        triangle, one query per lane for the rest.  stack_capacity: 0 keeps the compiled
        traversal stack, else 32, 64 or 128."""
        t = _BatchTuning(block_size, chunk, blocks_per_cu, traversal, stack_capacity, wave_queries_per_triangle,
                         0, 0)
        return _status(_lib.gwn_batch_set_tuning(self._h, ctypes.byref(t), _stream(stream)))

    def capture(self, include_refit: bool = True) -> Status:
        return _status(_lib.gwn_batch_capture(self._h, 1 if include_refit else 0))

    def replay(self, stream=None) -> Status:
        return _status(_lib.gwn_batch_replay(self._h, _stream(stream)))

    def status(self, stream=None) -> Status:
        """Synchronizes the stream; INTERNAL_ERROR if a refit met an invalid tree since the last call."""
        return _status(_lib.gwn_batch_status(self._h, _stream(stream)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.gwn_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BatchBuildItem(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_void_p), ("qx", ctypes.c_void_p), ("qy", ctypes.c_void_p),
                ("qz", ctypes.c_void_p), ("out", ctypes.c_void_p), ("count", ctypes.c_uint64)]


class _BatchBuildConfig(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("width", ctypes.c_int32), ("hploc_search_radius", ctypes.c_uint32),
                ("order", ctypes.c_int32)]


class BuiltBatchObject(BatchObject):
    """Batched mode over geometries (no reference counterpart): the batch builds and owns every
    mesh's tree (H-PLOC or LBVH) and Taylor moments in one segmented pass, which a captured graph replays
    together with the queries.  ``items`` are tuples (geometry, qx, qy, qz, out); each mesh's tree,
    moments and results equal a single-mesh build, refit and Taylor batch of that mesh."""

    def __init__(self, items, width: int = 4, hploc_radius: int = 8, order: int = 1, accuracy_scale: float = 2.0,
                 method: "BuildMethod" = None, stream=None):
        self._keep = list(items)
        arr = (_BatchBuildItem * max(1, len(self._keep)))()
        for k, (geo, qx, qy, qz, out) in enumerate(self._keep):
            n = _numel(qx)
            if _numel(qy) != n or _numel(qz) != n or _numel(out) != n:
                raise ValueError("Query SoA spans require identical lengths.")
            arr[k] = _BatchBuildItem(geo.handle.value, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(out), n)
        method = BuildMethod.HPLOC if method is None else method
        cfg = _BatchBuildConfig(int(method), int(width), int(hploc_radius), int(order))
        h = ctypes.c_void_p()
        _status(_lib.gwn_batch_create_built(arr, len(self._keep), ctypes.byref(cfg), float(accuracy_scale),
                                            ctypes.byref(h), _stream(stream))).raise_if_error()
        self._h = h

    def build(self, stream=None) -> Status:
        """Enqueues every mesh's build and moment refit."""
        return _status(_lib.gwn_batch_build(self._h, _stream(stream)))

    def set_build(self, width: int, hploc_radius: int, method: "BuildMethod" = None, stream=None) -> Status:
        method = BuildMethod.HPLOC if method is None else method
        cfg = _BatchBuildConfig(int(method), int(width), int(hploc_radius), 0)
        return _status(_lib.gwn_batch_set_build(self._h, ctypes.byref(cfg), _stream(stream)))

    def build_config(self) -> dict:
        cfg = _BatchBuildConfig()
        levels = ctypes.c_int32()
        _status(_lib.gwn_batch_get_build(self._h, ctypes.byref(cfg), ctypes.byref(levels))).raise_if_error()
        return {"method": BuildMethod(cfg.method), "width": cfg.width, "hploc_search_radius": cfg.hploc_search_radius,
                "order": cfg.order, "level_launches": levels.value}

    def tune_build(self, stream=None) -> float:
        ms = ctypes.c_float()
        _status(_lib.gwn_batch_tune_build(self._h, _stream(stream), ctypes.byref(ms))).raise_if_error()
        return float(ms.value)

    def export(self, mesh: int, stream=None):
        """(BvhObject, MomentObject) holding copies of mesh ``mesh``'s tree and moments."""
        bvh = BvhObject()
        mom = MomentObject(self.build_config()["order"])
        _status(_lib.gwn_batch_export(self._h, int(mesh), bvh.handle, mom.handle, _stream(stream))).raise_if_error()
        return bvh, mom


def gwn_compute_query_order(bvh: BvhObject, qx, qy, qz, order, stream=None) -> Status:
    """Hilbert visiting order of query points (uint32 device array, one entry per query)."""
    bad = _check_soa(qx, qy, qz, [order])
    if bad:
        return bad
    return _status(_lib.gwn_compute_query_order_f32(
        bvh.handle, _dptr(qx), _dptr(qy), _dptr(qz), _dptr(order), _numel(qx), _stream(stream)))


def radix_sort_pairs(keys, values, begin_bit: int = 0, end_bit: int | None = None, stream=None) -> Status:
    """Stable device sort of (uint64 or uint32 key, uint32 value) pairs in place."""
    import torch

    n = _numel(keys)
    if _numel(values) != n:
        return Status(Error.INVALID_ARGUMENT, "Key and value spans must have identical lengths.")
    if keys.dtype in (torch.int64, torch.uint64):
        fn, width = _lib.gwn_radix_sort_pairs_u64, 64
    else:
        fn, width = _lib.gwn_radix_sort_pairs_u32, 32
    end_bit = width if end_bit is None else end_bit
    return _status(fn(_dptr(keys), _dptr(values), n, int(begin_bit), int(end_bit), _stream(stream)))
