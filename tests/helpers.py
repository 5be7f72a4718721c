"""Independent checks shared by the oracle tests and the GPU parity tests.

Each check restates a property the mathematics or the reference's own tests fix,
written differently from the oracle (bit loops instead of magic-number spreads,
top-down recursion instead of Karras' parallel search, brute-force unions), so a
slip in the oracle or the HIP path fails here.
"""

from __future__ import annotations

import numpy as np

OFFSET_MASK = (1 << 47) - 1
VALID_BIT = 1 << 47

CHILD_DTYPE = np.dtype([("bounds", np.float32, (6,)), ("reference", np.uint64)])
NODE_DTYPE = np.dtype([("children", CHILD_DTYPE, (4,))])


def as_nodes(raw) -> np.ndarray:
    """View downloaded node bytes (or an oracle array) as the 128-byte BVH4 node dtype."""
    a = np.asarray(raw)
    if a.dtype == NODE_DTYPE:
        return a
    return np.frombuffer(np.ascontiguousarray(a).tobytes(), dtype=NODE_DTYPE)


def morton_bitloop_u64(qx: int, qy: int, qz: int) -> int:
    """Interleave 21-bit quantized axes bit by bit: x at 3b+2, y at 3b+1, z at 3b."""
    code = 0
    for b in range(21):
        code |= ((qx >> b) & 1) << (3 * b + 2)
        code |= ((qy >> b) & 1) << (3 * b + 1)
        code |= ((qz >> b) & 1) << (3 * b)
    return code


def morton_bitloop_u32(qx: int, qy: int, qz: int) -> int:
    code = 0
    for b in range(10):
        code |= ((qx >> b) & 1) << (3 * b + 2)
        code |= ((qy >> b) & 1) << (3 * b + 1)
        code |= ((qz >> b) & 1) << (3 * b)
    return code


def hilbert_key10(x, y, z) -> np.ndarray:
    """Position of 1024^3 grid cells along the 3-D Hilbert curve, Skilling's transpose method
    ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), vectorised: 30-bit keys with
    axis 0's transposed bit the most significant of each triple."""
    X = [np.asarray(v, dtype=np.uint64) & np.uint64(0x3FF) for v in (x, y, z)]
    Q = 1 << 9
    while Q > 1:
        P = np.uint64(Q - 1)
        q = np.uint64(Q)
        for i in range(3):
            bit = (X[i] & q) != 0
            t = (X[0] ^ X[i]) & P
            x0 = np.where(bit, X[0] ^ P, X[0] ^ t)
            if i:
                X[i] = np.where(bit, X[i], X[i] ^ t)
            X[0] = x0
        Q >>= 1
    X[1] = X[1] ^ X[0]
    X[2] = X[2] ^ X[1]
    t = np.zeros_like(X[0])
    Q = 1 << 9
    while Q > 1:
        t = np.where((X[2] & np.uint64(Q)) != 0, t ^ np.uint64(Q - 1), t)
        Q >>= 1
    X = [v ^ t for v in X]
    key = np.zeros_like(X[0])
    for b in range(10):
        for a in range(3):
            key |= ((X[a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 2 - a)
    return key


def query_cells(qx, qy, qz, root_bounds) -> tuple:
    """The presort's grid cell of each query: FP32 (q - min) * (1 / extent), clamped to [0, 1],
    times 1023, truncated (the library's arithmetic, operation by operation)."""
    lo = np.asarray(root_bounds[:3], np.float32)
    hi = np.asarray(root_bounds[3:], np.float32)
    out = []
    for k, q in enumerate((qx, qy, qz)):
        ext = np.float32(hi[k] - lo[k])
        s = np.float32(1.0) / ext if ext > 0 else np.float32(1.0)
        f = (np.asarray(q, np.float32) - lo[k]) * s
        f = np.minimum(np.maximum(f, np.float32(0.0)), np.float32(1.0)) * np.float32(1023.0)
        out.append(np.nan_to_num(f, nan=0.0).astype(np.uint32))
    return tuple(out)


def _delta(codes, i, j) -> int:
    """Common-prefix length of keys i and j, extended by the index for equal keys."""
    n = len(codes)
    if j < 0 or j >= n:
        return -1
    if codes[i] == codes[j]:
        x = i ^ j
        return 128 if x == 0 else 64 + (64 - x.bit_length())
    return 64 - int(codes[i] ^ codes[j]).bit_length()


def lbvh_topdown(codes) -> dict:
    """Sequential top-down radix-tree construction over sorted keys.

    Node ids follow Karras: the node covering [f, l] split at s has children s (covering
    [f, s]) and s + 1 (covering [s + 1, l]); leaves are sorted positions.  Returns
    {node_id: (left, left_is_leaf, right, right_is_leaf)}.
    """
    codes = [int(c) for c in codes]
    n = len(codes)
    out = {}
    if n < 2:
        return out
    todo = [(0, n - 1, 0)]
    while todo:
        f, l, nid = todo.pop()
        d = _delta(codes, f, l)
        s = f
        for c in range(f, l):
            if _delta(codes, f, c) > d:
                s = c
        left_leaf = s == f
        right_leaf = s + 1 == l
        out[nid] = (s, left_leaf, s + 1, right_leaf)
        if not left_leaf:
            todo.append((f, s, s))
        if not right_leaf:
            todo.append((s + 1, l, s + 1))
    return out


def half_area32(b) -> np.float32:
    b = np.asarray(b, dtype=np.float32)
    d = np.maximum(b[3:] - b[:3], np.float32(0))
    return np.float32(np.float32(d[0] * d[1]) + np.float32(d[1] * d[2])) + np.float32(d[2] * d[0])


def check_bvh4(nodes_raw, root_reference: int, primitive_indices, triangles, mesh,
               internal_stack_bound: int, packed_stack_bound: int, expect_bfs: bool = True):
    """Structural invariants of a canonical BVH4 (tests/integration_gwn_builder.cu:36-119)."""
    nodes = as_nodes(nodes_raw)
    T = len(primitive_indices)
    assert sorted(np.asarray(primitive_indices).tolist()) == list(range(T))
    # Records are the original vertices of each sorted primitive.
    v = np.stack([mesh.vx, mesh.vy, mesh.vz], axis=1).astype(np.float32)
    f = np.stack([mesh.i0, mesh.i1, mesh.i2], axis=1).astype(np.int64)[np.asarray(primitive_indices, dtype=np.int64)]
    expect_rec = np.concatenate([v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]], axis=1)
    assert np.array_equal(np.asarray(triangles, dtype=np.float32).view(np.uint32), expect_rec.view(np.uint32))
    tri_lo = np.minimum(np.minimum(expect_rec[:, 0:3], expect_rec[:, 3:6]), expect_rec[:, 6:9])
    tri_hi = np.maximum(np.maximum(expect_rec[:, 0:3], expect_rec[:, 3:6]), expect_rec[:, 6:9])
    if (root_reference >> 48) != 0:
        assert len(nodes) == 0 and internal_stack_bound == 0 and packed_stack_bound == 0
        return
    N = len(nodes)
    seen_leaf = np.zeros(T, dtype=np.int64)
    seen_node = np.zeros(N, dtype=np.int64)

    # Post-order recursion: subtree bounds, stack bounds, coverage.
    import sys

    sys.setrecursionlimit(100000)

    def rec(n, depth):
        assert 0 <= n < N
        seen_node[n] += 1
        lo = np.full(3, np.inf, dtype=np.float32)
        hi = np.full(3, -np.inf, dtype=np.float32)
        n_valid = n_int = 0
        best_int = best_packed = 0
        for s in range(4):
            ch = nodes[n]["children"][s]
            ref = int(ch["reference"])
            if not ref & VALID_BIT:
                continue
            n_valid += 1
            if (ref >> 48) == 0:
                c = ref & OFFSET_MASK
                if expect_bfs:
                    assert c > n, "breadth-first numbering puts children after parents"
                n_int += 1
                clo, chi, bi, bp = rec(c, depth + 1)
                best_int, best_packed = max(best_int, bi), max(best_packed, bp)
            else:
                off, cnt = ref & OFFSET_MASK, ref >> 48
                assert cnt == 1 and off < T
                seen_leaf[off] += 1
                clo, chi = tri_lo[off], tri_hi[off]
            b = np.asarray(ch["bounds"], dtype=np.float32)
            assert np.array_equal(b[:3].view(np.uint32), clo.view(np.uint32)), "child bounds = union of leaf boxes"
            assert np.array_equal(b[3:].view(np.uint32), chi.view(np.uint32))
            lo, hi = np.minimum(lo, clo), np.maximum(hi, chi)
        assert 2 <= n_valid <= 4
        si = (n_int - 1 + best_int) if n_int else 0
        sp = n_valid - 1 + best_packed
        return lo, hi, si, sp

    _, _, si, sp = rec(int(root_reference & OFFSET_MASK), 0)
    assert np.all(seen_leaf == 1), "every primitive is reached exactly once"
    assert np.all(seen_node == 1), "every node is reachable exactly once"
    assert si == internal_stack_bound
    assert sp == packed_stack_bound


def check_collapse_rule(nodes_raw, binary_nodes, binary_bounds, binary_root: int = 0, width: int = 4):
    """Largest-half-area expansion (include/gwn/detail/gwn_bvh_build_collapse.cuh:170-211).

    For every wide node, the binary nodes opened inside it all have half-area >= every
    internal binary node left on its frontier (ties broken toward the smaller index),
    and a node with fewer than `width` children has no internal frontier node left to open.
    Slots past the node's children (and past the width) are empty.
    """
    nodes = as_nodes(nodes_raw)
    bn = np.asarray(binary_nodes)
    areas = [half_area32(b) for b in np.asarray(binary_bounds)]
    # Map each binary internal node to the wide node that contains it.
    parent_of = {}
    for i in range(len(bn)):
        l, r, lk, rk = (int(x) for x in bn[i])
        if lk == 1:
            parent_of[l] = i
        if rk == 1:
            parent_of[r] = i
    # Frontier of each wide node: its internal children's binary roots are the binary nodes
    # whose bounds equal the child bounds; recover them by walking from the binary root.
    def frontier(broot):
        cands = [(int(bn[broot][0]), int(bn[broot][2])), (int(bn[broot][1]), int(bn[broot][3]))]
        opened = [broot]
        while len(cands) < width:
            best = None
            for k, (idx, kind) in enumerate(cands):
                if kind != 1:
                    continue
                if best is None or areas[idx] > areas[cands[best][0]] or (
                    areas[idx] == areas[cands[best][0]] and idx < cands[best][0]
                ):
                    best = k
            if best is None:
                break
            idx = cands[best][0]
            opened.append(idx)
            cands[best] = (int(bn[idx][0]), int(bn[idx][2]))
            cands.append((int(bn[idx][1]), int(bn[idx][3])))
        return opened, cands

    # Walk wide nodes from the binary root and compare child order/bounds.
    stack = [(0, int(binary_root))]  # (wide node, binary root)
    while stack:
        w, b = stack.pop()
        opened, cands = frontier(b)
        inner = [areas[o] for o in opened[1:]]
        for idx, kind in cands:
            if kind == 1 and inner:
                assert areas[idx] <= min(inner)
        kids = nodes[w]["children"]
        for s, (idx, kind) in enumerate(cands):
            ref = int(kids[s]["reference"])
            assert ref & VALID_BIT
            if kind == 2:
                assert (ref >> 48) == 1 and (ref & OFFSET_MASK) == idx
            else:
                assert (ref >> 48) == 0
                assert np.array_equal(np.asarray(kids[s]["bounds"], np.float32).view(np.uint32),
                                      np.asarray(binary_bounds[idx], np.float32).view(np.uint32))
                stack.append((ref & OFFSET_MASK, idx))
        for s in range(len(cands), 4):
            assert not int(kids[s]["reference"]) & VALID_BIT


def triangle_second_moment(a, b, c) -> np.ndarray:
    """Closed form: int over the triangle of (p - P)(p - P)^T dA = A/12 * sum_k d_k d_k^T, P = centroid."""
    a, b, c = (np.asarray(x, dtype=np.float64) for x in (a, b, c))
    P = (a + b + c) / 3.0
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    M = np.zeros((3, 3))
    for v in (a, b, c):
        d = v - P
        M += np.outer(d, d)
    return area / 12.0 * M


def solid_angle_quad(a, b, c, q, n: int = 64) -> float:
    """Solid angle by Gauss-Legendre quadrature of (p - q).n / |p - q|^3 over the triangle."""
    a, b, c, q = (np.asarray(x, dtype=np.float64) for x in (a, b, c, q))
    x, w = np.polynomial.legendre.leggauss(n)
    u = 0.5 * (x + 1.0)
    wu = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    WU, WV = np.meshgrid(wu, wu, indexing="ij")
    # Duffy map of the unit square onto the triangle: s = U, t = U * V, Jacobian U.
    s = U
    t = U * V
    p = a + s[..., None] * (b - a) + t[..., None] * (c - b)
    nvec = np.cross(b - a, c - a)
    r = p - q
    integrand = (r @ nvec) / np.linalg.norm(r, axis=-1) ** 3
    return float(np.sum(integrand * U * WU * WV))
