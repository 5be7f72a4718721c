"""PLY triangle-mesh reader: ASCII and binary (little- and big-endian) files.

Replaces the reference's mesh input (tests/test_mesh_io.cpp:43-92, which calls libigl's
read_triangle_mesh): vertex positions x, y, z of any numeric PLY type become float32 SoA arrays,
and the face element's vertex list becomes three uint32 index arrays.  A face of k > 3 vertices
is split into the fan (v0, v[i], v[i+1]), i = 1 .. k-2.  Other elements and properties are read
past and ignored.

Like the reference's loader, ``load_ply_mesh`` returns None (instead of raising) for a file that
does not describe a usable triangle mesh: no vertices or no triangles, a non-finite coordinate,
an index outside the vertex range, or a vertex count above the uint32 index range.  Malformed
files (bad header, truncated body) raise ``PlyError``.

``collect_ply_model_paths`` lists a directory's ``.ply`` files in sorted order
(tests/test_mesh_io.cpp:20-31).
"""

from __future__ import annotations

import os
import sys

import numpy as np

_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


class PlyError(ValueError):
    """The file is not a readable PLY file."""


class _Property:
    def __init__(self, name: str, dtype: str, count_dtype: str | None = None):
        self.name = name
        self.dtype = dtype              # scalar type, or the list's item type
        self.count_dtype = count_dtype  # list length type; None for a scalar property

    @property
    def is_list(self) -> bool:
        return self.count_dtype is not None


class _Element:
    def __init__(self, name: str, count: int):
        self.name = name
        self.count = count
        self.props: list[_Property] = []


def _parse_header(f) -> tuple[str, list[_Element]]:
    if f.readline().strip() != b"ply":
        raise PlyError("missing 'ply' magic line")
    fmt = None
    elements: list[_Element] = []
    while True:
        line = f.readline()
        if not line:
            raise PlyError("header ends before 'end_header'")
        words = line.decode("ascii", "replace").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        key = words[0]
        if key == "format":
            if len(words) < 3 or words[1] not in ("ascii", "binary_little_endian", "binary_big_endian"):
                raise PlyError(f"unsupported format line: {line!r}")
            fmt = words[1]
        elif key == "element":
            if len(words) != 3:
                raise PlyError(f"bad element line: {line!r}")
            try:
                count = int(words[2])
            except ValueError as exc:
                raise PlyError(f"bad element count: {line!r}") from exc
            if count < 0:
                raise PlyError(f"negative element count: {line!r}")
            elements.append(_Element(words[1], count))
        elif key == "property":
            if not elements:
                raise PlyError("property before any element")
            if len(words) == 5 and words[1] == "list":
                if words[2] not in _TYPES or words[3] not in _TYPES:
                    raise PlyError(f"unknown list types: {line!r}")
                elements[-1].props.append(_Property(words[4], _TYPES[words[3]], _TYPES[words[2]]))
            elif len(words) == 3 and words[1] in _TYPES:
                elements[-1].props.append(_Property(words[2], _TYPES[words[1]]))
            else:
                raise PlyError(f"bad property line: {line!r}")
        elif key == "end_header":
            break
        else:
            raise PlyError(f"unknown header keyword: {line!r}")
    if fmt is None:
        raise PlyError("missing format line")
    return fmt, elements


def _read_ascii(f, elements):
    tokens = f.read().split()
    pos = 0
    data = {}
    for el in elements:
        cols: dict[str, list] = {p.name: [] for p in el.props}
        for _ in range(el.count):
            for p in el.props:
                if p.is_list:
                    if pos >= len(tokens):
                        raise PlyError(f"element '{el.name}' truncated")
                    k = int(tokens[pos])
                    pos += 1
                    if k < 0 or pos + k > len(tokens):
                        raise PlyError(f"element '{el.name}' truncated")
                    cols[p.name].append([int(t) for t in tokens[pos:pos + k]])
                    pos += k
                else:
                    if pos >= len(tokens):
                        raise PlyError(f"element '{el.name}' truncated")
                    cols[p.name].append(float(tokens[pos]) if p.dtype[0] == "f" else int(tokens[pos]))
                    pos += 1
        data[el.name] = cols
    return data


def _read_binary(f, elements, endian: str):
    buf = f.read()
    pos = 0
    data = {}
    for el in elements:
        if not any(p.is_list for p in el.props):
            dt = np.dtype([(p.name, endian + p.dtype) for p in el.props])
            nbytes = dt.itemsize * el.count
            if pos + nbytes > len(buf):
                raise PlyError(f"element '{el.name}' truncated")
            rec = np.frombuffer(buf, dtype=dt, count=el.count, offset=pos)
            pos += nbytes
            data[el.name] = {p.name: rec[p.name] for p in el.props}
            continue
        # Variable-length records.  Fast path: one list property whose every length is the same
        # (all triangles, say), so the element is a fixed-stride table; verified, else parsed one
        # record at a time.
        fixed = _try_fixed_lists(buf, pos, el, endian)
        if fixed is not None:
            data[el.name], pos = fixed
            continue
        cols: dict[str, list] = {p.name: [] for p in el.props}
        for _ in range(el.count):
            for p in el.props:
                if p.is_list:
                    cdt = np.dtype(endian + p.count_dtype)
                    if pos + cdt.itemsize > len(buf):
                        raise PlyError(f"element '{el.name}' truncated")
                    k = int(np.frombuffer(buf, cdt, 1, pos)[0])
                    pos += cdt.itemsize
                    idt = np.dtype(endian + p.dtype)
                    if k < 0 or pos + k * idt.itemsize > len(buf):
                        raise PlyError(f"element '{el.name}' truncated")
                    cols[p.name].append(np.frombuffer(buf, idt, k, pos).astype(np.int64).tolist())
                    pos += k * idt.itemsize
                else:
                    sdt = np.dtype(endian + p.dtype)
                    if pos + sdt.itemsize > len(buf):
                        raise PlyError(f"element '{el.name}' truncated")
                    cols[p.name].append(np.frombuffer(buf, sdt, 1, pos)[0].item())
                    pos += sdt.itemsize
        data[el.name] = cols
    return data


def _try_fixed_lists(buf, pos, el, endian):
    lists = [p for p in el.props if p.is_list]
    if len(lists) != 1 or el.count == 0:
        return None
    lp = lists[0]
    # Length of the first record's list decides the candidate stride.
    off = pos
    for p in el.props:
        if p is lp:
            break
        off += np.dtype(p.dtype).itemsize
    cdt = np.dtype(endian + lp.count_dtype)
    if off + cdt.itemsize > len(buf):
        return None
    k = int(np.frombuffer(buf, cdt, 1, off)[0])
    fields = []
    for p in el.props:
        if p is lp:
            fields.append(("__count", endian + p.count_dtype))
            fields.append((p.name, endian + p.dtype, (k,)) if k > 0 else (p.name, endian + p.dtype, (0,)))
        else:
            fields.append((p.name, endian + p.dtype))
    dt = np.dtype(fields)
    nbytes = dt.itemsize * el.count
    if pos + nbytes > len(buf):
        return None
    rec = np.frombuffer(buf, dtype=dt, count=el.count, offset=pos)
    if not np.all(rec["__count"] == k):
        return None
    cols = {p.name: (rec[p.name] if not p.is_list else rec[p.name].astype(np.int64)) for p in el.props}
    return cols, pos + nbytes


def read_ply(path: str) -> dict:
    """Every element of the file as {element: {property: values}}; list properties are an
    (n, k) integer array when every list has length k, else a list of lists."""
    with open(path, "rb") as f:
        fmt, elements = _parse_header(f)
        if fmt == "ascii":
            try:
                return _read_ascii(f, elements)
            except ValueError as exc:
                if isinstance(exc, PlyError):
                    raise
                raise PlyError(f"bad ASCII value: {exc}") from exc
        return _read_binary(f, elements, "<" if fmt == "binary_little_endian" else ">")


def _triangles(face_lists) -> np.ndarray:
    """(T, 3) int64 fan triangulation of the face lists (faces of fewer than 3 vertices dropped)."""
    if isinstance(face_lists, np.ndarray):
        k = face_lists.shape[1] if face_lists.ndim == 2 else 0
        if k < 3:
            return np.zeros((0, 3), np.int64)
        parts = [np.stack([face_lists[:, 0], face_lists[:, i], face_lists[:, i + 1]], 1) for i in range(1, k - 1)]
        # Keep each face's fan triangles consecutive, in face order.
        return np.stack(parts, 1).reshape(-1, 3).astype(np.int64)
    tris = []
    for face in face_lists:
        for i in range(1, len(face) - 1):
            tris.append((face[0], face[i], face[i + 1]))
    return np.asarray(tris, dtype=np.int64).reshape(-1, 3)


def load_ply_mesh(path: str):
    """A ``synth.Mesh`` (float32 positions, uint32 indices) from a PLY file, or None when the file
    is not a ``.ply`` path or holds no usable triangle mesh (tests/test_mesh_io.cpp:43-92)."""
    from smallgwn_amd import synth

    if not str(path).endswith(".ply"):
        return None
    data = read_ply(path)
    vert = data.get("vertex")
    face = data.get("face")
    if vert is None or face is None or not all(k in vert for k in ("x", "y", "z")):
        return None
    lists = face.get("vertex_indices", face.get("vertex_index"))
    if lists is None:
        return None
    xyz = np.stack([np.asarray(vert[k], dtype=np.float64) for k in ("x", "y", "z")], 1)
    tris = _triangles(lists)
    nv = xyz.shape[0]
    if nv == 0 or tris.shape[0] == 0 or nv > 0xFFFFFFFF:
        return None
    if not np.isfinite(xyz).all():
        return None
    if (tris < 0).any() or (tris >= nv).any():
        return None
    return synth._mesh_from_arrays(xyz.astype(np.float32), tris.astype(np.uint32))


def collect_ply_model_paths(model_dir: str) -> list[str]:
    """Sorted ``.ply`` regular files of a directory (tests/test_mesh_io.cpp:20-31)."""
    out = []
    for name in os.listdir(model_dir):
        p = os.path.join(model_dir, name)
        if os.path.isfile(p) and name.endswith(".ply"):
            out.append(p)
    return sorted(out)


def write_ply(path: str, mesh, binary: bool = True) -> None:
    """Writes a mesh as PLY (float x, y, z; uchar/uint vertex_indices lists), for round trips."""
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], 1).astype(np.float32)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], 1).astype(np.uint32)
    fmt = "binary_little_endian" if binary else "ascii"
    header = (f"ply\nformat {fmt} 1.0\nelement vertex {v.shape[0]}\nproperty float x\nproperty float y\n"
              f"property float z\nelement face {f.shape[0]}\nproperty list uchar uint vertex_indices\nend_header\n")
    with open(path, "wb") as out:
        out.write(header.encode("ascii"))
        if binary:
            out.write(v.astype("<f4").tobytes())
            rec = np.zeros(f.shape[0], dtype=[("n", "u1"), ("i", "<u4", (3,))])
            rec["n"] = 3
            rec["i"] = f
            out.write(rec.tobytes())
        else:
            for row in v:
                out.write(f"{float(row[0])!r} {float(row[1])!r} {float(row[2])!r}\n".encode("ascii"))
            for row in f:
                out.write(f"3 {int(row[0])} {int(row[1])} {int(row[2])}\n".encode("ascii"))


if __name__ == "__main__":
    for p in sys.argv[1:]:
        m = load_ply_mesh(p)
        print(p, "unusable" if m is None else f"{m.vertex_count} vertices, {m.triangle_count} triangles")
