"""Seeded synthetic meshes, query sets and ray sets for the geometric-query path.

Every random stream is a splitmix64 sequence, so the same seed produces the same
float32 inputs on any host (no implementation-defined distribution objects).
A uniform variate is ``(x >> 40) * 2**-24``: 24 random bits, exactly
representable in float32.

The default seed ``0xB3A9E5D4`` is the one the reference benchmarks use
(tests/benchmark_main.cu:37).  The mesh families follow the workloads the
reference's benchmarks and tests exercise: a subdivided icosphere, radially
displaced latitude-longitude spheres, open triangle soups and tori.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

DEFAULT_SEED = 0xB3A9E5D4

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX1 = np.uint64(0xBF58476D1CE4E5B9)
_MIX2 = np.uint64(0x94D049BB133111EB)
_STREAM = 0xD1B54A32D192ED03
_MASK64 = (1 << 64) - 1


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _MIX1
    z = (z ^ (z >> np.uint64(27))) * _MIX2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, n: int) -> np.ndarray:
    """Return the first ``n`` outputs of splitmix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & _MASK64) + i * _GOLDEN
        return _mix(z)


def substream(seed: int, stream_id: int) -> int:
    """Derive an independent 64-bit seed for stream ``stream_id``."""
    z = (seed + (stream_id + 1) * _STREAM) & _MASK64
    with np.errstate(over="ignore"):
        return int(_mix(np.array([z], dtype=np.uint64))[0])


def uniform01(seed: int, n: int) -> np.ndarray:
    """``n`` float64 variates in [0, 1) carrying 24 random bits each."""
    bits = splitmix64(seed, n) >> np.uint64(40)
    return bits.astype(np.float64) * (1.0 / (1 << 24))


@dataclass
class Mesh:
    """Triangle mesh in the SoA layout the library uploads (float32 xyz, uint32 indices)."""

    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    i0: np.ndarray
    i1: np.ndarray
    i2: np.ndarray

    @property
    def vertex_count(self) -> int:
        return int(self.vx.shape[0])

    @property
    def triangle_count(self) -> int:
        return int(self.i0.shape[0])

    def vertices(self) -> np.ndarray:
        return np.stack([self.vx, self.vy, self.vz], axis=1)

    def faces(self) -> np.ndarray:
        return np.stack([self.i0, self.i1, self.i2], axis=1)


def _mesh_from_arrays(v: np.ndarray, f: np.ndarray) -> Mesh:
    v32 = np.ascontiguousarray(v, dtype=np.float32)
    f32 = np.ascontiguousarray(f, dtype=np.uint32)
    return Mesh(
        np.ascontiguousarray(v32[:, 0]),
        np.ascontiguousarray(v32[:, 1]),
        np.ascontiguousarray(v32[:, 2]),
        np.ascontiguousarray(f32[:, 0]),
        np.ascontiguousarray(f32[:, 1]),
        np.ascontiguousarray(f32[:, 2]),
    )


def icosphere(subdivisions: int = 4, radius: float = 1.0) -> Mesh:
    """Closed, outward-oriented icosphere: 10*4^s+2 vertices, 20*4^s triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        cache: dict[tuple[int, int], int] = {}

        def midpoint(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                idx = len(v) - 1
                cache[key] = idx
            return idx

        new_faces = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    vv = np.array(v) * radius
    return _mesh_from_arrays(vv, np.array(faces, dtype=np.int64))


def _real_sph_harm(l: int, m: int, theta: np.ndarray, phi: np.ndarray, phase: float) -> np.ndarray:
    """Orthonormal real spherical harmonic of degree l, order m >= 0, with a phase shift."""
    from scipy.special import lpmv

    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
    if m > 0:
        norm *= math.sqrt(2.0)
    return norm * lpmv(m, l, np.cos(theta)) * np.cos(m * phi + phase)


def harmonic_radius_terms(seed: int, count: int = 8, max_degree: int = 12):
    """Seeded (degree, order, phase) triples of the radial displacement field."""
    u = uniform01(substream(seed, 101), 3 * count).reshape(count, 3)
    terms = []
    for k in range(count):
        l = 1 + min(int(u[k, 0] * max_degree), max_degree - 1)
        m = min(int(u[k, 1] * (l + 1)), l)
        terms.append((l, m, 2.0 * math.pi * u[k, 2]))
    return terms


def displaced_sphere(
    nlon: int, nlat: int, seed: int = DEFAULT_SEED, amplitude: float = 0.05, harmonics: int = 8
) -> Mesh:
    """Latitude-longitude sphere, r = 1 + amplitude * sum of random-phase real harmonics.

    Two triangles per quad over ``nlon x nlat`` quads, so the triangle count is
    exactly ``2 * nlon * nlat`` (1024 x 512 gives 1,048,576).  The two poles are
    single vertices; each pole quad yields one zero-area triangle with a
    repeated vertex index, which exercises the zero-area paths of moment refit.
    """
    terms = harmonic_radius_terms(seed, harmonics)
    ring_theta = math.pi * np.arange(1, nlat) / nlat
    ring_phi = 2.0 * math.pi * np.arange(nlon) / nlon
    theta = np.concatenate([[0.0], np.repeat(ring_theta, nlon), [math.pi]])
    phi = np.concatenate([[0.0], np.tile(ring_phi, nlat - 1), [0.0]])
    r = np.ones_like(theta)
    for l, m, ph in terms:
        r += amplitude * _real_sph_harm(l, m, theta, phi, ph)
    st = np.sin(theta)
    v = np.stack([r * st * np.cos(phi), r * st * np.sin(phi), r * np.cos(theta)], axis=1)

    south = 1 + (nlat - 1) * nlon

    def vid(i: np.ndarray, j: np.ndarray) -> np.ndarray:
        jj = j % nlon
        out = 1 + (i - 1) * nlon + jj
        out = np.where(i == 0, 0, out)
        return np.where(i == nlat, south, out)

    ii, jj = np.meshgrid(np.arange(nlat), np.arange(nlon), indexing="ij")
    ii = ii.ravel()
    jj = jj.ravel()
    a = vid(ii, jj)
    b = vid(ii + 1, jj)
    c = vid(ii + 1, jj + 1)
    d = vid(ii, jj + 1)
    f = np.empty((2 * ii.size, 3), dtype=np.int64)
    f[0::2] = np.stack([a, b, c], axis=1)
    f[1::2] = np.stack([a, c, d], axis=1)
    return _mesh_from_arrays(v, f)


def torus(nu: int, nv: int, major: float = 1.0, minor: float = 0.35) -> Mesh:
    """Closed, outward-oriented torus around the z axis with 2*nu*nv triangles."""
    u = 2.0 * math.pi * np.arange(nu) / nu
    w = 2.0 * math.pi * np.arange(nv) / nv
    uu, ww = np.meshgrid(u, w, indexing="ij")
    rr = major + minor * np.cos(ww)
    v = np.stack([rr * np.cos(uu), rr * np.sin(uu), minor * np.sin(ww)], axis=-1).reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = ii * nv + jj
    b = ((ii + 1) % nu) * nv + jj
    c = ((ii + 1) % nu) * nv + (jj + 1) % nv
    d = ii * nv + (jj + 1) % nv
    f = np.empty((2 * a.size, 3), dtype=np.int64)
    f[0::2] = np.stack([a.ravel(), b.ravel(), c.ravel()], axis=1)
    f[1::2] = np.stack([a.ravel(), c.ravel(), d.ravel()], axis=1)
    return _mesh_from_arrays(v, f)


def _face_frames(mesh: Mesh):
    v = mesh.vertices().astype(np.float64)
    f = mesh.faces().astype(np.int64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    length = np.linalg.norm(n, axis=1)
    return a, b, c, n, length


def surface_points(mesh: Mesh, seed: int, n: int):
    """Uniform-by-index points on non-degenerate faces with their unit face normals."""
    a, b, c, nrm, length = _face_frames(mesh)
    good = np.nonzero(length > 0)[0]
    u = uniform01(seed, 3 * n).reshape(n, 3)
    face = good[np.minimum((u[:, 0] * good.size).astype(np.int64), good.size - 1)]
    s = np.sqrt(u[:, 1])
    b0 = 1.0 - s
    b1 = s * (1.0 - u[:, 2])
    b2 = s * u[:, 2]
    p = b0[:, None] * a[face] + b1[:, None] * b[face] + b2[:, None] * c[face]
    unit = nrm[face] / length[face, None]
    return p, unit


def _soa32(p: np.ndarray):
    p32 = p.astype(np.float32)
    return (
        np.ascontiguousarray(p32[:, 0]),
        np.ascontiguousarray(p32[:, 1]),
        np.ascontiguousarray(p32[:, 2]),
    )


def uniform_box_points(seed: int, n: int, lo, hi):
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    u = uniform01(seed, 3 * n).reshape(n, 3)
    return _soa32(lo + u * (hi - lo))


def mixed_queries(mesh: Mesh, n: int, seed: int = DEFAULT_SEED, near_fraction: float = 0.5,
                  box: float = 1.5, band: float = 0.01):
    """Queries: (1-near_fraction) uniform in [-box, box]^3, the rest on-surface +- U[-band, band] along the normal."""
    n_near = int(round(n * near_fraction))
    n_far = n - n_near
    ux, uy, uz = uniform_box_points(substream(seed, 1), n_far, (-box,) * 3, (box,) * 3)
    p, unit = surface_points(mesh, substream(seed, 2), n_near)
    off = (2.0 * uniform01(substream(seed, 3), n_near) - 1.0) * band
    nx, ny, nz = _soa32(p + off[:, None] * unit)
    return (
        np.concatenate([ux, nx]),
        np.concatenate([uy, ny]),
        np.concatenate([uz, nz]),
    )


def near_surface_log_queries(mesh: Mesh, n: int, seed: int = DEFAULT_SEED,
                             lo: float = 1e-6, hi: float = 1e-1):
    """Surface point + normal * s * (+-1), s log-uniform in [lo, hi]."""
    p, unit = surface_points(mesh, substream(seed, 4), n)
    u = uniform01(substream(seed, 5), 2 * n).reshape(n, 2)
    s = np.exp(math.log(lo) + u[:, 0] * (math.log(hi) - math.log(lo)))
    sign = np.where(u[:, 1] < 0.5, -1.0, 1.0)
    return _soa32(p + (sign * s)[:, None] * unit)


def rays_toward_ball(n: int, seed: int = DEFAULT_SEED, origin_radius: float = 3.0):
    """Origins uniform on a sphere of ``origin_radius``; unit directions toward uniform points in the unit ball."""
    u = uniform01(substream(seed, 6), 5 * n).reshape(n, 5)
    z = 2.0 * u[:, 0] - 1.0
    phi = 2.0 * math.pi * u[:, 1]
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    o = origin_radius * np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], axis=1)
    z2 = 2.0 * u[:, 2] - 1.0
    phi2 = 2.0 * math.pi * u[:, 3]
    rxy2 = np.sqrt(np.maximum(0.0, 1.0 - z2 * z2))
    t = np.cbrt(u[:, 4])[:, None] * np.stack([rxy2 * np.cos(phi2), rxy2 * np.sin(phi2), z2], axis=1)
    d = t - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ox, oy, oz = _soa32(o)
    dx, dy, dz = _soa32(d)
    return ox, oy, oz, dx, dy, dz


def open_soup(nlon: int, nlat: int, seed: int = DEFAULT_SEED, delete_fraction: float = 0.30,
              jitter_fraction: float = 0.02, jitter_scale: float = 1e-4) -> Mesh:
    """Displaced sphere with a seeded fraction of faces deleted and a fraction given jittered duplicate vertices."""
    base = displaced_sphere(nlon, nlat, seed)
    v = base.vertices().astype(np.float64)
    f = base.faces().astype(np.int64)
    u = uniform01(substream(seed, 7), f.shape[0])
    f = f[u >= delete_fraction]
    e = np.concatenate([
        np.linalg.norm(v[f[:, 1]] - v[f[:, 0]], axis=1),
        np.linalg.norm(v[f[:, 2]] - v[f[:, 1]], axis=1),
        np.linalg.norm(v[f[:, 0]] - v[f[:, 2]], axis=1),
    ])
    sigma = jitter_scale * float(e.mean())
    pick = uniform01(substream(seed, 8), f.shape[0]) < jitter_fraction
    idx = np.nonzero(pick)[0]
    g = uniform01(substream(seed, 9), 6 * idx.size).reshape(idx.size * 3, 2)
    gauss = np.sqrt(-2.0 * np.log(np.maximum(g[:, 0], 1e-300))) * np.cos(2.0 * math.pi * g[:, 1])
    gauss3 = np.stack([gauss, np.roll(gauss, 1), np.roll(gauss, 2)], axis=1) * sigma
    new_v = v[f[idx].ravel()] + gauss3
    new_ids = v.shape[0] + np.arange(idx.size * 3)
    f[idx] = new_ids.reshape(idx.size, 3)
    v = np.concatenate([v, new_v], axis=0)
    return _mesh_from_arrays(v, f)


# Small analytic meshes used by the reference's unit tests (tests/test_meshes.hpp:13-53,
# tests/unit_gwn_query_winding.cu:23-30).

def octahedron() -> Mesh:
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    return _mesh_from_arrays(v, f)


def asymmetric_octahedron() -> Mesh:
    x = np.array([1.7, -0.8, 0.15, -0.1, 0.2, -0.2], dtype=np.float32)
    y = np.array([0.1, -0.2, 1.2, -0.7, 0.1, -0.1], dtype=np.float32)
    z = np.array([0.0, 0.1, -0.1, 0.15, 1.5, -0.9], dtype=np.float32)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    return _mesh_from_arrays(np.stack([x, y, z], axis=1), f)


def closed_cube() -> Mesh:
    vx = [-1, 1, 1, -1, -1, 1, 1, -1]
    vy = [-1, -1, 1, 1, -1, -1, 1, 1]
    vz = [-1, -1, -1, -1, 1, 1, 1, 1]
    i0 = [0, 0, 4, 4, 0, 0, 1, 1, 0, 0, 3, 3]
    i1 = [2, 3, 5, 6, 4, 7, 2, 6, 1, 5, 7, 6]
    i2 = [1, 2, 6, 7, 7, 3, 6, 5, 5, 4, 6, 2]
    return _mesh_from_arrays(np.stack([vx, vy, vz], axis=1), np.stack([i0, i1, i2], axis=1))


def layered_triangles(count: int) -> Mesh:
    """`count` copies of the triangle (0,0,z), (1,0,z), (0,1,z) at z = 0, 1, ..., count - 1
    (the layered fixture of the reference's distance and ray tests)."""
    v = []
    f = []
    for k in range(count):
        v += [[0.0, 0.0, float(k)], [1.0, 0.0, float(k)], [0.0, 1.0, float(k)]]
        f.append([3 * k, 3 * k + 1, 3 * k + 2])
    return _mesh_from_arrays(np.asarray(v, dtype=np.float32), np.asarray(f, dtype=np.int64))


def open_cube() -> Mesh:
    vx = [-1, 1, 1, -1, -1, 1, 1, -1]
    vy = [-1, -1, 1, 1, -1, -1, 1, 1]
    vz = [-1, -1, -1, -1, 1, 1, 1, 1]
    i0 = [0, 0, 0, 0, 1, 1, 0, 0, 3, 3]
    i1 = [2, 3, 4, 7, 2, 6, 1, 5, 7, 6]
    i2 = [1, 2, 7, 3, 6, 5, 5, 4, 6, 2]
    return _mesh_from_arrays(np.stack([vx, vy, vz], axis=1), np.stack([i0, i1, i2], axis=1))


# Workload shapes of the benchmarked configurations.

def config_small(seed: int = DEFAULT_SEED):
    """Icosphere subdivided 4x (2562 verts, 5120 tris); 65,536 queries in [-1.5,1.5]^3; 65,536 rays."""
    mesh = icosphere(4)
    qx, qy, qz = uniform_box_points(substream(seed, 1), 65536, (-1.5,) * 3, (1.5,) * 3)
    return mesh, (qx, qy, qz), rays_toward_ball(65536, seed)


def config_large_mesh(seed: int = DEFAULT_SEED) -> Mesh:
    """Radially displaced sphere with 1,048,576 triangles (1024 x 512 lat-long quads)."""
    return displaced_sphere(1024, 512, seed)
