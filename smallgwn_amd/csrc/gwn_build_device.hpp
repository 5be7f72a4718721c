// gwn_build_device.hpp - device pieces of the BVH4 build shared by the single-tree build
// (gwn_build.hip) and the segmented build of many trees in one pass (gwn_batch_build.hip).
//
// Every piece that touches a tree takes that tree's place in the arrays it reads (its first
// sorted leaf and its first binary node, `build_segment`); the single-tree build passes the
// segment {0, T, 0}.  Both builds therefore run the same arithmetic in the same order and
// produce the same trees bit for bit.
//
// Included by .hip translation units only.  The collapse compares half-areas and the Morton path
// compares quantized centroids: every float expression stays unfused (contract(off) below, in
// force for the including file from here on), so the CPU oracle reproduces them bit for bit.
#pragma once

#include "gwn_core.hpp"
#include "gwn_lookback.hpp"

#include <cmath>

#pragma clang fp contract(off)

namespace gwn {
namespace {

constexpr int k_block = 256;
constexpr std::uint32_t k_invalid = 0xffffffffu;
constexpr std::uint32_t k_leaf_bit = 0x80000000u;

__device__ inline float fmin_ref(float a, float b) { return (b < a) ? b : a; } // std::min(a, b)
__device__ inline float fmax_ref(float a, float b) { return (a < b) ? b : a; } // std::max(a, b)

__device__ inline std::uint32_t float_order_key(float f) {
    std::uint32_t const u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float float_from_order_key(std::uint32_t k) {
    std::uint32_t const u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

// One tree's place in the build arrays: sorted leaves [leaf_base, leaf_base + leaves) and binary
// internal nodes [internal_base, internal_base + leaves - 1).  Child references stored in the
// binary nodes are local to the tree (leaf and node indices counted from these bases).
struct build_segment {
    std::uint32_t leaf_base;
    std::uint32_t leaves;
    std::uint32_t internal_base;
    std::uint32_t tree; // index of the tree (its root word)
};

// Per-axis min/max over vertices i, i + stride, ... < nv; one atomic per block per extremum into
// the tree's six order-key words.  The whole block calls it.
__device__ inline void scene_bounds_range(
    float const *__restrict__ x, float const *__restrict__ y, float const *__restrict__ z, std::size_t i,
    std::size_t nv, std::size_t stride, std::uint32_t *__restrict__ ordered
) {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (; i < nv; i += stride) {
        float const p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], p[k]);
            hi[k] = fmaxf(hi[k], p[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int d = 32; d >= 1; d >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], d));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], d));
        }
    }
    // One atomic per bound and block (the waves meet in LDS first): same-address atomics from
    // every wave of every block had serialised.
    __shared__ float s_b[k_block / 64][6];
    int const wave = static_cast<int>(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s_b[wave][k] = lo[k];
            s_b[wave][3 + k] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 1; w < k_block / 64; ++w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo[k] = fminf(lo[k], s_b[w][k]);
                hi[k] = fmaxf(hi[k], s_b[w][3 + k]);
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            atomicMin(&ordered[k], float_order_key(lo[k]));
            atomicMax(&ordered[3 + k], float_order_key(hi[k]));
        }
    }
}

__device__ inline void scene_init_words(std::uint32_t *ordered) {
    for (int k = 0; k < 3; ++k) {
        ordered[k] = ~0u;
        ordered[3 + k] = 0u;
    }
}

// scene[0..8] = min xyz, max xyz, inverse extent xyz (the reference's host safe_inv).
__device__ inline void scene_finalize_words(std::uint32_t const *ordered, float *scene) {
    for (int k = 0; k < 3; ++k) {
        float const lo = float_from_order_key(ordered[k]);
        float const hi = float_from_order_key(ordered[3 + k]);
        scene[k] = lo;
        scene[3 + k] = hi;
        scene[6 + k] = (hi == lo) ? 1.0f : 1.0f / (hi - lo);
    }
}

__device__ inline std::uint64_t expand_bits_u64(std::uint64_t v) {
    std::uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ inline float clamp01(float v) {
    if (v < 0.0f)
        return 0.0f;
    if (v > 1.0f)
        return 1.0f;
    return v;
}

// AABB and 63-bit Morton code of triangle t's centroid within the scene box (scene: min xyz,
// max xyz, inverse extent xyz).  A triangle with an out-of-range vertex gets a zero box and code
// (the records pass reports it).
__device__ inline void morton_triangle(
    float const *__restrict__ vx, float const *__restrict__ vy, float const *__restrict__ vz, std::size_t nv,
    std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1, std::uint32_t const *__restrict__ i2,
    std::size_t t, float const *__restrict__ scene, aabb &box, std::uint64_t &code
) {
    std::uint32_t const a = i0[t], b = i1[t], c = i2[t];
    if (a >= nv || b >= nv || c >= nv) {
        box = aabb{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        code = 0;
        return;
    }
    float const ax = vx[a], ay = vy[a], az = vz[a];
    float const bx = vx[b], by = vy[b], bz = vz[b];
    float const cx = vx[c], cy = vy[c], cz = vz[c];
    box = aabb{
        fmin_ref(ax, fmin_ref(bx, cx)), fmin_ref(ay, fmin_ref(by, cy)), fmin_ref(az, fmin_ref(bz, cz)),
        fmax_ref(ax, fmax_ref(bx, cx)), fmax_ref(ay, fmax_ref(by, cy)), fmax_ref(az, fmax_ref(bz, cz)),
    };
    float const mx = (box.min_x + box.max_x) * 0.5f;
    float const my = (box.min_y + box.max_y) * 0.5f;
    float const mz = (box.min_z + box.max_z) * 0.5f;
    float const nx = (mx - scene[0]) * scene[6];
    float const ny = (my - scene[1]) * scene[7];
    float const nz = (mz - scene[2]) * scene[8];
    std::uint64_t const qx = static_cast<std::uint64_t>(clamp01(nx) * 2097151.0f);
    std::uint64_t const qy = static_cast<std::uint64_t>(clamp01(ny) * 2097151.0f);
    std::uint64_t const qz = static_cast<std::uint64_t>(clamp01(nz) * 2097151.0f);
    code = (expand_bits_u64(qx) << 2) | (expand_bits_u64(qy) << 1) | expand_bits_u64(qz);
}

// Leaf-ordered triangle record of primitive p; false (nothing written) for an out-of-range vertex.
__device__ inline bool triangle_record(
    float const *__restrict__ vx, float const *__restrict__ vy, float const *__restrict__ vz, std::size_t nv,
    std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1, std::uint32_t const *__restrict__ i2,
    std::uint32_t p, tri_record *__restrict__ out
) {
    std::uint32_t const a = i0[p], b = i1[p], c = i2[p];
    if (a >= nv || b >= nv || c >= nv)
        return false;
    *out = tri_record{vx[a], vy[a], vz[a], vx[b], vy[b], vz[b], vx[c], vy[c], vz[c]};
    return true;
}

__device__ inline aabb aabb_union_ref(aabb const &a, aabb const &b) {
    return aabb{fmin_ref(a.min_x, b.min_x), fmin_ref(a.min_y, b.min_y), fmin_ref(a.min_z, b.min_z),
                fmax_ref(a.max_x, b.max_x), fmax_ref(a.max_y, b.max_y), fmax_ref(a.max_z, b.max_z)};
}

__device__ inline float half_area(aabb const &b) {
    float dx = b.max_x - b.min_x;
    float dy = b.max_y - b.min_y;
    float dz = b.max_z - b.min_z;
    dx = dx < 0.0f ? 0.0f : dx;
    dy = dy < 0.0f ? 0.0f : dy;
    dz = dz < 0.0f ? 0.0f : dz;
    return dx * dy + dy * dz + dz * dx;
}

// delta(i, j) of Karras' construction with the reference's equal-code tie-break
// (gwn_bvh_build_binary.cuh:37-57).
__device__ inline int lbvh_delta(std::uint64_t const *codes, std::int64_t n, std::int64_t i, std::int64_t j) {
    if (j < 0 || j >= n)
        return -1;
    std::uint64_t const ci = codes[i];
    std::uint64_t const cj = codes[j];
    if (ci == cj) {
        std::uint64_t const diff = static_cast<std::uint64_t>(i) ^ static_cast<std::uint64_t>(j);
        return diff == 0 ? 128 : 64 + __clzll(static_cast<long long>(diff));
    }
    return __clzll(static_cast<long long>(ci ^ cj));
}

// Internal node i of the Karras LBVH over the n sorted codes of one tree (tree-local indices;
// `codes`, `bnodes`, `iparent` and `lparent` at the tree's bases).  Binary node = (left, right) with
// bit 31 marking a leaf child.  Parent links carry the side in bit 0: parent = (node << 1) | side.
// lparent may be null.
__device__ inline void lbvh_node(
    std::uint64_t const *__restrict__ codes, std::int64_t N, std::int64_t i, uint2 *__restrict__ bnodes,
    std::uint32_t *__restrict__ iparent, std::uint32_t *__restrict__ lparent
) {
    int const dir = (lbvh_delta(codes, N, i, i + 1) - lbvh_delta(codes, N, i, i - 1) >= 0) ? 1 : -1;
    int const dmin = lbvh_delta(codes, N, i, i - dir);
    std::int64_t lmax = 2;
    while (lbvh_delta(codes, N, i, i + lmax * dir) > dmin)
        lmax <<= 1;
    std::int64_t len = 0;
    for (std::int64_t step = lmax >> 1; step > 0; step >>= 1)
        if (lbvh_delta(codes, N, i, i + (len + step) * dir) > dmin)
            len += step;
    std::int64_t const j = i + len * dir;
    std::int64_t const first = i < j ? i : j;
    std::int64_t const last = i < j ? j : i;
    int const dnode = lbvh_delta(codes, N, first, last);
    std::int64_t split = first;
    std::int64_t step = last - first;
    do {
        step = (step + 1) >> 1;
        std::int64_t const cand = split + step;
        if (cand < last && lbvh_delta(codes, N, first, cand) > dnode)
            split = cand;
    } while (step > 1);
    bool const left_leaf = split == first;
    bool const right_leaf = split + 1 == last;
    std::uint32_t const l = static_cast<std::uint32_t>(split);
    std::uint32_t const r = static_cast<std::uint32_t>(split + 1);
    bnodes[i] = make_uint2(l | (left_leaf ? k_leaf_bit : 0u), r | (right_leaf ? k_leaf_bit : 0u));
    std::uint32_t const me = static_cast<std::uint32_t>(i) << 1;
    if (!left_leaf)
        iparent[l] = me;
    else if (lparent)
        lparent[l] = me;
    if (!right_leaf)
        iparent[r] = me | 1u;
    else if (lparent)
        lparent[r] = me | 1u;
    if (i == 0)
        iparent[0] = k_invalid;
}


// Depth of internal node i (edges to the root) by following the parent links of its tree; a walk
// longer than the node count means a broken parent chain: error, depth 0.
__device__ inline std::uint32_t lbvh_node_depth(std::uint32_t const *__restrict__ iparent, std::size_t ni,
                                                std::size_t i, std::uint32_t *error) {
    std::uint32_t d = 0;
    std::uint32_t link = iparent[i];
    while (link != k_invalid) {
        if (++d > ni || (link >> 1) >= ni) {
            atomicOr(error, 1u);
            return 0;
        }
        link = iparent[link >> 1];
    }
    return d;
}

// Bounds and first leaf of internal node `node` from its children (leaves, or nodes of a deeper
// level written by an earlier launch), unioned in left/right order (the reference's order).
__device__ inline void lbvh_node_bounds(aabb const *__restrict__ leaf_boxes, uint2 const *__restrict__ bnodes,
                                        std::size_t node, aabb *__restrict__ ibounds,
                                        std::uint32_t *__restrict__ ifirst) {
    uint2 const bn = bnodes[node];
    std::uint32_t const l = bn.x & ~k_leaf_bit, r = bn.y & ~k_leaf_bit;
    aabb const lb = (bn.x & k_leaf_bit) ? leaf_boxes[l] : ibounds[l];
    aabb const rb = (bn.y & k_leaf_bit) ? leaf_boxes[r] : ibounds[r];
    std::uint32_t const lf = (bn.x & k_leaf_bit) ? l : ifirst[l];
    std::uint32_t const rf = (bn.y & k_leaf_bit) ? r : ifirst[r];
    ibounds[node] = aabb_union_ref(lb, rb);
    ifirst[node] = lf < rf ? lf : rf;
}

// H-PLOC binary build (reference: include/gwn/detail/gwn_bvh_build_hploc.cuh:104-622), wave64.
//
// Every sorted leaf is a lane that climbs the implicit LBVH of the (code, position) keys: at each
// step it exchanges its range end into the boundary it meets (hploc.cuh:141-156, 451-470); the
// first arrival stops, the second joins both ranges.  A joined range with more than 16 leaves, or
// the full range, concatenates its children's cluster lists (at most 16 each, stored at the start
// of each child range) and runs PLOC on them: nearest neighbour within the search radius by the
// packed key (union half-area bits << 1, low four bits = (offset - 1) << 1 | direction parity,
// hploc.cuh:195-241), mutual pairs merged into the lower lane, survivors compacted in lane order,
// until at most 16 clusters remain (one for the full range); the list goes back to the start of
// the range.  The 16-cluster lists are the reference's, so the merge sequence of every range is
// the reference's; a 64-lane wave runs two ranges at once, one per 32-lane half.
//
// Node ids do not come from an atomic counter: merging clusters whose smallest sorted leaves are
// a < b creates the tree's internal node b - 1 (counted from its first leaf).  Each leaf is the
// larger-side minimum of at most one merge, so ids are unique, cover [0, T - 2], and do not depend
// on scheduling; bounds and first leaves are written beside the node, so no bounds climb follows.
// Release before the boundary exchange and acquire after it publish a range's list and nodes to
// the lane (on any XCD) that joins it.
//
// Trees side by side (the segmented build): a tree's first and last leaves are hard range ends,
// equal codes break ties by tree-local positions, and a range covering the whole tree is final.
// Cluster ids are global (a leaf: its position; an internal node: total_leaves + its global
// index); binary nodes store tree-local child references.
constexpr std::uint32_t k_hploc_list = 16; // clusters kept per partial range (hploc.cuh:23)
constexpr int k_hploc_block = 256;

__device__ inline std::uint64_t hploc_delta(std::uint64_t const *codes, std::uint32_t a, std::uint32_t b,
                                            std::uint32_t base) {
    std::uint64_t const d = codes[a] ^ codes[b];
    return d == 0 ? static_cast<std::uint64_t>((a - base) ^ (b - base)) : d;
}

__device__ inline std::uint32_t hploc_area_key(aabb const &a, aabb const &b) {
    float const area = half_area(aabb_union_ref(a, b));
    return (__float_as_uint(area) << 1) & ~15u;
}

// Lane-local view of one half wave: bits of a 64-lane ballot that belong to this lane's half.
__device__ inline std::uint32_t half_bits(std::uint64_t m) {
    return static_cast<std::uint32_t>(m >> (threadIdx.x & 32u));
}

__device__ inline aabb shfl_box32(aabb const &b, std::uint32_t src) {
    int const s = static_cast<int>(src);
    return aabb{__shfl(b.min_x, s, 32), __shfl(b.min_y, s, 32), __shfl(b.min_z, s, 32),
                __shfl(b.max_x, s, 32), __shfl(b.max_y, s, 32), __shfl(b.max_z, s, 32)};
}

// Position of the k-th (0-based) set bit of `mask`; requires k < popcount(mask).
__device__ inline std::uint32_t select_bit(std::uint32_t mask, std::uint32_t k) {
    std::uint32_t pos = 0;
#pragma unroll
    for (std::uint32_t w = 16; w > 0; w >>= 1) {
        std::uint32_t const low = __popc(mask & ((1u << w) - 1u));
        if (k >= low) {
            k -= low;
            mask >>= w;
            pos += w;
        }
    }
    return pos;
}

struct hploc_out {
    uint2 *bnodes;
    aabb *ibounds;
    std::uint32_t *ifirst; // global first sorted leaf of each node
    std::uint32_t *clusters;
};

// Tree-local reference of cluster id c of tree `seg` (leaf bit for a leaf).
__device__ inline std::uint32_t hploc_local_ref(std::uint32_t c, std::uint32_t total_leaves, build_segment const &seg) {
    return c < total_leaves ? ((c - seg.leaf_base) | k_leaf_bit) : (c - total_leaves - seg.internal_base);
}

// PLOC over one range [L, R1) split at S of tree `seg`, run by the 32 lanes of this half (all
// active, uniform control flow within the half).  Returns the merges made; sets `bad` on an
// inconsistent state.
__device__ inline std::uint32_t hploc_range(
    aabb const *__restrict__ leaf_boxes, std::uint32_t total_leaves, std::uint32_t radius, hploc_out const &o,
    build_segment const &seg, std::uint32_t L, std::uint32_t R1, std::uint32_t S, bool fin,
    std::uint32_t *root_words, bool &bad
) {
    std::uint32_t const h = threadIdx.x & 31u;
    std::uint32_t const node_limit = seg.leaves - 1u; // tree-local internal ids are below this
    // Concatenate the two child lists (hploc.cuh:159-177, 424-430).
    std::uint32_t idx = k_invalid;
    std::uint32_t const lcl = (S - L) < k_hploc_list ? (S - L) : k_hploc_list;
    if (h < lcl)
        idx = o.clusters[L + h];
    std::uint32_t const num_left = __popc(half_bits(__ballot(h < lcl && idx != k_invalid)));
    std::uint32_t const lcr = (R1 - S) < k_hploc_list ? (R1 - S) : k_hploc_list;
    std::uint32_t const ri = h - num_left;
    bool const rv = ri < lcr;
    if (rv)
        idx = o.clusters[S + ri];
    std::uint32_t const num_right = __popc(half_bits(__ballot(rv && idx != k_invalid)));
    std::uint32_t const total = num_left + num_right;
    std::uint32_t num = total;
    std::uint32_t first = k_invalid;
    aabb box{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (h < num) {
        if (idx < total_leaves) {
            first = idx;
            box = leaf_boxes[idx];
        } else if (idx - total_leaves - seg.internal_base < node_limit) {
            first = o.ifirst[idx - total_leaves];
            box = o.ibounds[idx - total_leaves];
        } else {
            bad = true;
        }
    }
    std::uint32_t const threshold = fin ? 1u : k_hploc_list;
    std::uint32_t made = 0;
    while (num > threshold) {
        // Smallest packed key over the pairs (h, h + r) and (h - r, h), r <= radius.
        std::uint32_t best = 0xffffffffu;
#pragma unroll
        for (std::uint32_t r = 1; r <= 8u; ++r) {
            if (r > radius)
                break;
            bool const nv = h + r < num;
            aabb const nb = shfl_box32(box, nv ? h + r : h);
            std::uint32_t ak = 0xffffffffu;
            if (nv) {
                ak = hploc_area_key(box, nb);
                std::uint32_t const key = ak | ((r - 1u) << 1) | (h & 1u);
                best = key < best ? key : best;
            }
            std::uint32_t const from_left = __shfl(ak, static_cast<int>(h >= r ? h - r : h), 32);
            if (h >= r && h < num) {
                std::uint32_t const key = from_left | ((r - 1u) << 1) | ((h & 1u) ^ 1u);
                best = key < best ? key : best;
            }
        }
        std::uint32_t const off = ((best & 15u) >> 1) + 1u;
        std::uint32_t const nn = (((best ^ h) & 1u) == 0u) ? h + off : h - off;
        // Mutual nearest neighbours merge; the lower lane keeps the cluster (hploc.cuh:337-404).
        bool const act = h < num;
        bool const nvalid = nn < num && nn != h;
        if (act && !nvalid)
            bad = true;
        std::uint32_t const safe = nvalid ? nn : h;
        std::uint32_t const back = __shfl(safe, static_cast<int>(safe), 32);
        bool const mutual = act && nvalid && back == h;
        bool const merge = mutual && h < safe;
        std::uint32_t const n_idx = __shfl(idx, static_cast<int>(safe), 32);
        std::uint32_t const n_first = __shfl(first, static_cast<int>(safe), 32);
        aabb const n_box = shfl_box32(box, safe);
        if (merge) {
            std::uint32_t const id = (first > n_first ? first : n_first) - seg.leaf_base - 1u;
            if (first == n_first || id >= node_limit) {
                bad = true;
            } else {
                aabb const merged = aabb_union_ref(box, n_box);
                std::uint32_t const f = first < n_first ? first : n_first;
                std::uint32_t const g = seg.internal_base + id;
                o.bnodes[g] = make_uint2(hploc_local_ref(idx, total_leaves, seg),
                                         hploc_local_ref(n_idx, total_leaves, seg));
                o.ibounds[g] = merged;
                o.ifirst[g] = f;
                idx = total_leaves + g;
                first = f;
                box = merged;
            }
        }
        std::uint32_t const merges = __popc(half_bits(__ballot(merge)));
        std::uint32_t const keep_mask = half_bits(__ballot(merge || !mutual));
        std::uint32_t const kept = __popc(keep_mask);
        std::uint32_t const src = h < kept ? select_bit(keep_mask, h) : h;
        idx = __shfl(idx, static_cast<int>(src), 32);
        first = __shfl(first, static_cast<int>(src), 32);
        box = shfl_box32(box, src);
        if (h >= kept)
            idx = k_invalid;
        num -= merges;
        made += merges;
        if (merges == 0) {
            bad = true;
            break;
        }
    }
    if (h < total)
        o.clusters[L + h] = idx;
    if (fin && h == 0u && num == 1u)
        root_words[seg.tree] = idx - total_leaves - seg.internal_base;
    return made;
}

// H-PLOC runs in two launches with the same decisions and merges as one global climb:
//   1. chunk climb: a workgroup takes the sorted leaves of its 1024-leaf chunk, one per lane,
//      and climbs while the boundary a range meets lies between two leaves of the chunk: those
//      boundaries live in LDS and workgroup-scope release / acquire order the ranges' lists and
//      nodes (no L2 write-back).  A range whose next step meets a boundary outside the chunk
//      parks; only a range holding the chunk's first or last leaf can, so at most two per chunk.
//      At the end the chunk's boundary words go to the global array.
//   2. global climb: the parked ranges go on from where they stopped over the global boundary
//      array, with agent-scope ordering as the one-launch climb has.  For one large tree every
//      parked range is large and merges at nearly every step, so each gets a 32-lane half of its
//      own (lane 0 of the half climbs): a wave then runs its two ranges' PLOC side by side instead
//      of one after the other for each of up to 64 ranges.  Many small trees park many short
//      ranges: one per lane keeps the launch small.
// Which lane arrives first at a boundary can differ from one launch to the other, but each
// boundary is met once from each side and the second arrival joins the same two ranges, so every
// merge is the same.  A range that covers its whole tree stops after its final PLOC.
// Chunk lengths: one tree, and many trees side by side.  With one parked range per lane in the
// global launch, 4,096 meshes of 113M triangles build (with their moments) in 116 / 102 / 101 ms
// at 256 / 512 / 1024-leaf chunks.
constexpr int k_hploc_chunk_block = 1024;
constexpr int k_hploc_seg_chunk_block = 1024;

// Parked range of lane `threadIdx.x` of block `block` in the global launch (k_hploc_block
// threads per block, one range per 32-lane half), or k_invalid for the other lanes.
__device__ inline std::uint32_t hploc_parked_slot(std::uint32_t block) {
    if ((threadIdx.x & 31u) != 0u)
        return k_invalid;
    return block * (static_cast<std::uint32_t>(k_hploc_block) / 32u) + (threadIdx.x >> 5);
}
// Blocks of the global launch for `parked` ranges.
constexpr std::size_t hploc_parked_blocks(std::size_t parked) {
    return (parked + k_hploc_block / 32 - 1) / (k_hploc_block / 32);
}
// Most ranges a chunk launch can park: two per chunk of `chunk` leaves.
constexpr std::size_t hploc_parked_capacity(std::size_t leaves, std::size_t chunk) {
    return 2 * ((leaves + chunk - 1) / chunk);
}

__device__ inline void release_workgroup() {
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}
__device__ inline void acquire_workgroup() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

// The climb of the ranges [left, right] of tree `seg` held by `active` lanes; every lane of the
// wave calls it.  Chunk: boundaries in LDS (`boundary` indexed from chunk_lo), parking ranges
// that leave [chunk_lo, chunk_hi) into `parked`; otherwise the global boundary array.  words:
// [0] merges, [1] failure flag; root_words[tree] receives each tree's tree-local root.
template <bool Chunk>
__device__ inline void hploc_climb(
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t total_leaves,
    std::uint32_t radius, hploc_out const &o, std::uint32_t *boundary, build_segment seg, std::uint32_t left,
    std::uint32_t right, bool active, std::uint32_t chunk_lo, std::uint32_t chunk_hi, uint2 *parked,
    std::uint32_t *parked_count, std::uint32_t *root_words, std::uint32_t *words
) {
    std::uint32_t const lane = threadIdx.x & 63u;
    std::uint32_t split = 0;
    std::uint32_t const last = seg.leaf_base + seg.leaves - 1u;
    bool bad = false;
    std::uint32_t made = 0;
    for (;;) {
        if constexpr (Chunk)
            release_workgroup();
        else
            release_agent();
        if (active) {
            bool const to_right =
                left == seg.leaf_base ||
                (right != last && hploc_delta(codes, right, right + 1u, seg.leaf_base) <
                                      hploc_delta(codes, left - 1u, left, seg.leaf_base));
            std::uint32_t const slot = to_right ? right : left - 1u;
            bool park = false;
            if constexpr (Chunk)
                park = slot < chunk_lo || slot + 1u >= chunk_hi;
            if (park) {
                parked[atomicAdd(parked_count, 1u)] = make_uint2(left, right);
                active = false;
            } else {
                std::uint32_t const want = to_right ? left : right;
                std::uint32_t prev;
                if constexpr (Chunk)
                    prev = __hip_atomic_exchange(&boundary[slot - chunk_lo], want, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
                else
                    prev = __hip_atomic_exchange(&boundary[slot], want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (prev == k_invalid) {
                    active = false;
                } else if (to_right) {
                    split = right + 1u;
                    right = prev;
                } else {
                    split = left;
                    left = prev;
                }
            }
        }
        if constexpr (Chunk)
            acquire_workgroup();
        else
            acquire_agent();
        bool const fin = active && right - left + 1u == seg.leaves;
        std::uint64_t pending = __ballot(active && (right - left + 1u > k_hploc_list || fin));
        while (pending != 0u) {
            int const s0 = __builtin_ctzll(pending);
            pending &= pending - 1u;
            int s1 = -1;
            if (pending != 0u) {
                s1 = __builtin_ctzll(pending);
                pending &= pending - 1u;
            }
            int const sel = (lane & 32u) ? s1 : s0;
            int const src = sel >= 0 ? sel : s0;
            std::uint32_t const rl = __shfl(left, src, 64);
            std::uint32_t const rr = __shfl(right, src, 64);
            std::uint32_t const rs = __shfl(split, src, 64);
            int const rf = __shfl(fin ? 1 : 0, src, 64);
            build_segment const rseg{__shfl(seg.leaf_base, src, 64), __shfl(seg.leaves, src, 64),
                                     __shfl(seg.internal_base, src, 64), __shfl(seg.tree, src, 64)};
            if (sel >= 0)
                made += hploc_range(leaf_boxes, total_leaves, radius, o, rseg, rl, rr + 1u, rs, rf != 0, root_words,
                                    bad);
        }
        if (fin)
            active = false; // the tree is complete
        if (__ballot(active) == 0u)
            break;
    }
    // `made` is the same on the 32 lanes of a half: count it once per half.
    std::uint32_t const total = wave_sum_u32((lane & 31u) == 0u ? made : 0u);
    bool const any_bad = __ballot(bad) != 0u;
    if (lane == 0u) {
        atomicAdd(&words[0], total);
        if (any_bad)
            atomicOr(&words[1], 1u);
    }
}

// The chunk launch's body for the block's chunk of sorted leaves [chunk_lo, chunk_hi) (Chunk =
// the block size): boundary words in LDS, then written back to the global array for the global
// launch.  `seg` / `active` are this lane's leaf's tree and whether it takes part (a tree of at
// least two leaves).
template <std::uint32_t Chunk>
__device__ inline void hploc_chunk_climb(
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t total_leaves,
    std::uint32_t radius, hploc_out const &o, std::uint32_t *boundary, build_segment seg, std::uint32_t leaf,
    bool active, uint2 *parked, std::uint32_t *parked_count, std::uint32_t *root_words, std::uint32_t *words
) {
    __shared__ std::uint32_t s_boundary[Chunk];
    std::uint32_t const chunk_lo = blockIdx.x * Chunk;
    std::uint32_t const chunk_hi = total_leaves - chunk_lo < Chunk ? total_leaves : chunk_lo + Chunk;
    s_boundary[threadIdx.x] = k_invalid;
    __syncthreads();
    hploc_climb<true>(codes, leaf_boxes, total_leaves, radius, o, s_boundary, seg, leaf, leaf, active, chunk_lo,
                      chunk_hi, parked, parked_count, root_words, words);
    __syncthreads();
    // Boundaries between two leaves of the chunk (the chunk's last slot faces the next chunk).
    if (chunk_lo + threadIdx.x + 1u < chunk_hi)
        boundary[chunk_lo + threadIdx.x] = s_boundary[threadIdx.x];
}

// Level-synchronous BVH4 collapse.  Level d holds the wide nodes at depth d, numbered
// contiguously from levels[d].start in (depth, first sorted leaf) order.  One launch per level:
// every thread takes one wide node of the level, expands its binary subtree to at most four
// candidates by always opening the internal candidate with the largest half-area (ties: smaller
// binary index; collapse.cuh:170-211), and a look-back scan over the level's internal-child
// counts numbers the next level's nodes.  Children of one node are numbered in first-leaf
// order, which keeps the whole level in first-leaf order, so nodes come out in their final
// breadth-first positions and no reordering pass follows.
// Wide depth never exceeds binary depth: at most 128 levels for LBVH (one (code, index) prefix bit
// per level); H-PLOC trees are deeper only for adversarial inputs, which fail the build with an
// internal error past this limit.
constexpr int k_max_levels = 1024;

struct collapse_task {
    std::uint32_t binary;
    std::uint32_t prefix_internal; // internal-child stack demand accumulated above this node
    std::uint32_t prefix_packed;   // same for all valid children
    std::uint32_t pad;
};

struct collapse_level {
    std::uint32_t start;
    std::uint32_t count;
    std::uint32_t tiles_claimed;
    std::uint32_t pad;
};

struct collapse_stats {
    std::uint32_t max_internal;
    std::uint32_t max_packed;
    std::uint32_t leaves;
    std::uint32_t error;
};

// One tree's collapse arrays, every pointer already at the tree's base (binary nodes hold
// tree-local references, so the binary arrays start at the tree's first binary node and the leaf
// boxes at its first sorted leaf).
struct collapse_tree {
    uint2 const *bnodes;
    aabb const *ibounds;
    std::uint32_t const *ifirst;
    aabb const *leaf_boxes;
    collapse_level *levels;
    collapse_task *tasks;
    collapse_stats *stats;
    bvh_node4 *nodes;
    std::uint32_t *parent;
    std::uint8_t *internal_children;
    std::uint32_t capacity;
};

__device__ inline std::uint32_t wave_max_u32(std::uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        std::uint32_t const w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// Tile `tile` (k_block consecutive wide nodes) of level `depth`, holding nodes [start, start +
// count) of tree `t`; `status` is the tree's look-back array for this level.  The whole block
// calls it with the same arguments; the caller has checked start + count <= capacity and that
// the tile lies within the level.
__device__ inline void collapse_tile(
    collapse_tree const &t, int depth, std::uint32_t start, std::uint32_t count, std::uint32_t tile,
    std::uint32_t *status, int width
) {
    __shared__ std::uint32_t s_exclusive;
    __shared__ std::uint32_t s_wave[k_block / 64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    std::uint32_t const i = tile * k_block + threadIdx.x;
    bool const active = i < count;
    std::uint32_t const self = start + i;

    std::uint32_t cand[4] = {k_invalid, k_invalid, k_invalid, k_invalid};
    int n_cand = 0;
    collapse_task task{};
    if (active) {
        task = t.tasks[self];
        uint2 const root = t.bnodes[task.binary];
        cand[0] = root.x;
        cand[1] = root.y;
        n_cand = 2;
        // Expand up to the tree's width (2, 3 or 4 children; slots past it stay empty).
        while (n_cand < width) {
            int best = -1;
            float best_area = -INFINITY;
            std::uint32_t best_index = k_invalid;
            for (int s = 0; s < n_cand; ++s) {
                if (cand[s] & k_leaf_bit)
                    continue;
                float const area = half_area(t.ibounds[cand[s]]);
                if (best >= 0 && area < best_area)
                    continue;
                if (best >= 0 && area == best_area && cand[s] >= best_index)
                    continue;
                best = s;
                best_area = area;
                best_index = cand[s];
            }
            if (best < 0)
                break;
            uint2 const ex = t.bnodes[cand[best]];
            cand[best] = ex.x;
            cand[n_cand] = ex.y;
            ++n_cand;
        }
    }
    std::uint32_t first[4];
    std::uint32_t n_internal = 0, n_leaf = 0;
    for (int s = 0; s < 4; ++s) {
        first[s] = k_invalid;
        if (s >= n_cand)
            continue;
        if (cand[s] & k_leaf_bit) {
            ++n_leaf;
        } else {
            ++n_internal;
            first[s] = t.ifirst[cand[s]];
        }
    }

    // Block-wide exclusive scan of internal-child counts, then the cross-tile look-back.
    std::uint32_t incl = n_internal;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        std::uint32_t const v = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += v;
    }
    if (lane == 63)
        s_wave[wave] = incl;
    __syncthreads();
    std::uint32_t wave_prefix = 0, block_total = 0;
    for (int w = 0; w < k_block / 64; ++w) {
        if (w < wave)
            wave_prefix += s_wave[w];
        block_total += s_wave[w];
    }
    if (wave == 0) {
        std::uint32_t const ex = wave_lookback(status, tile, block_total, &t.stats->error);
        if (lane == 0)
            s_exclusive = ex;
    }
    __syncthreads();
    std::uint32_t const next_start = start + count;
    if (threadIdx.x == 0 && tile == (count - 1) / k_block) {
        t.levels[depth + 1].start = next_start;
        t.levels[depth + 1].count = s_exclusive + block_total;
    }
    if (active) {
        std::uint32_t const base = next_start + s_exclusive + wave_prefix + incl - n_internal;
        std::uint32_t const node_internal = task.prefix_internal + (n_internal ? n_internal - 1u : 0u);
        std::uint32_t const node_packed = task.prefix_packed + static_cast<std::uint32_t>(n_cand - 1);
        bvh_node4 out{};
        for (int s = 0; s < n_cand; ++s) {
            std::uint32_t const c = cand[s] & ~k_leaf_bit;
            if (cand[s] & k_leaf_bit) {
                out.children[s].bounds = t.leaf_boxes[c];
                out.children[s].reference = encode_leaf(c, 1u);
                continue;
            }
            std::uint32_t rank = 0;
            for (int u = 0; u < 4; ++u)
                rank += first[u] < first[s] ? 1u : 0u;
            std::uint32_t const id = base + rank;
            if (id >= t.capacity) {
                atomicOr(&t.stats->error, 1u);
                continue;
            }
            out.children[s].bounds = t.ibounds[c];
            out.children[s].reference = encode_internal(id);
            t.tasks[id] = collapse_task{c, node_internal, node_packed, 0u};
            t.parent[id] = (self << 2) | static_cast<std::uint32_t>(s);
        }
        t.nodes[self] = out;
        t.internal_children[self] = static_cast<std::uint8_t>(n_internal);
        task.prefix_internal = node_internal;
        task.prefix_packed = node_packed;
    }
    std::uint32_t const w_int = wave_max_u32(active ? task.prefix_internal : 0u);
    std::uint32_t const w_pack = wave_max_u32(active ? task.prefix_packed : 0u);
    std::uint32_t const w_leaf = wave_sum_u32(n_leaf);
    if (lane == 0) {
        atomicMax(&t.stats->max_internal, w_int);
        atomicMax(&t.stats->max_packed, w_pack);
        atomicAdd(&t.stats->leaves, w_leaf);
    }
}

} // namespace
} // namespace gwn
