// gwn_build.hip - canonical BVH4 construction on gfx950.
//
// Pipeline (reference: include/gwn/detail/gwn_bvh_build_impl.cuh:44-255):
//   scene bounds (device reduction, no host read-back)       build_common.cuh:243-274
//   triangle AABB + 63-bit Morton code                       build_common.cuh:157-214
//   stable LSD radix sort of (code, triangle id)             build_common.cuh:325-347
//   Karras LBVH                                              build_binary.cuh:29-129
//   bottom-up bounds + first leaf (arrival counters)         build_binary.cuh:131-192
//   BVH4 collapse, one launch per level, emitting nodes      build_collapse.cuh:98-314
//     directly in breadth-first (depth, first sorted leaf)   (replaces reorder.cuh:61-143)
//     order
//   leaf-ordered triangle records                            build_impl.cuh:23-41
// The host reads the small level table back once per chunk of levels (about three times for a
// million triangles) to find the last level and size the node array; the reference performs
// one such read-back of its node count (collapse.cuh:425-446).
#include "gwn_core.hpp"
#include "gwn_lookback.hpp"

#include <algorithm>
#include <vector>

// The collapse compares half-areas and the Morton path compares quantized centroids; keep
// every float expression unfused so the CPU oracle reproduces them bit for bit.
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

// Per-axis min/max over all vertices; one atomic per block per extremum.
__global__ __launch_bounds__(k_block) void scene_bounds_kernel(
    float const *__restrict__ x, float const *__restrict__ y, float const *__restrict__ z,
    std::size_t nv, std::uint32_t *__restrict__ ordered
) {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::size_t i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    std::size_t const stride = static_cast<std::size_t>(gridDim.x) * k_block;
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

__global__ void scene_init_kernel(std::uint32_t *ordered) {
    for (int k = 0; k < 3; ++k) {
        ordered[k] = ~0u;
        ordered[3 + k] = 0u;
    }
}

// scene[0..8] = min xyz, max xyz, inverse extent xyz (the reference's host safe_inv).
__global__ void scene_finalize_kernel(std::uint32_t const *ordered, float *scene) {
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

__global__ __launch_bounds__(k_block) void morton_kernel(
    float const *__restrict__ vx, float const *__restrict__ vy, float const *__restrict__ vz,
    std::size_t nv, std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, std::size_t nt, float const *__restrict__ scene,
    aabb *__restrict__ boxes, std::uint64_t *__restrict__ codes, std::uint32_t *__restrict__ ids
) {
    std::size_t const t = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (t >= nt)
        return;
    ids[t] = static_cast<std::uint32_t>(t);
    std::uint32_t const a = i0[t], b = i1[t], c = i2[t];
    if (a >= nv || b >= nv || c >= nv) {
        boxes[t] = aabb{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        codes[t] = 0;
        return;
    }
    float const ax = vx[a], ay = vy[a], az = vz[a];
    float const bx = vx[b], by = vy[b], bz = vz[b];
    float const cx = vx[c], cy = vy[c], cz = vz[c];
    aabb const box{
        fmin_ref(ax, fmin_ref(bx, cx)), fmin_ref(ay, fmin_ref(by, cy)), fmin_ref(az, fmin_ref(bz, cz)),
        fmax_ref(ax, fmax_ref(bx, cx)), fmax_ref(ay, fmax_ref(by, cy)), fmax_ref(az, fmax_ref(bz, cz)),
    };
    boxes[t] = box;
    float const mx = (box.min_x + box.max_x) * 0.5f;
    float const my = (box.min_y + box.max_y) * 0.5f;
    float const mz = (box.min_z + box.max_z) * 0.5f;
    float const nx = (mx - scene[0]) * scene[6];
    float const ny = (my - scene[1]) * scene[7];
    float const nz = (mz - scene[2]) * scene[8];
    std::uint64_t const qx = static_cast<std::uint64_t>(clamp01(nx) * 2097151.0f);
    std::uint64_t const qy = static_cast<std::uint64_t>(clamp01(ny) * 2097151.0f);
    std::uint64_t const qz = static_cast<std::uint64_t>(clamp01(nz) * 2097151.0f);
    codes[t] = (expand_bits_u64(qx) << 2) | (expand_bits_u64(qy) << 1) | expand_bits_u64(qz);
}

__global__ __launch_bounds__(k_block) void gather_boxes_kernel(
    aabb const *__restrict__ boxes, std::uint32_t const *__restrict__ order, std::size_t nt,
    aabb *__restrict__ sorted
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < nt)
        sorted[i] = boxes[order[i]];
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

// Binary node = (left, right) with bit 31 marking a leaf child.  Parent links carry the side in
// bit 0: parent = (node << 1) | side.
__global__ __launch_bounds__(k_block) void lbvh_kernel(
    std::uint64_t const *__restrict__ codes, std::size_t n, uint2 *__restrict__ bnodes,
    std::uint32_t *__restrict__ iparent, std::uint32_t *__restrict__ lparent
) {
    std::size_t const tid = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (tid + 1 >= n)
        return;
    std::int64_t const i = static_cast<std::int64_t>(tid);
    std::int64_t const N = static_cast<std::int64_t>(n);
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
    if (left_leaf)
        lparent[l] = me;
    else
        iparent[l] = me;
    if (right_leaf)
        lparent[r] = me | 1u;
    else
        iparent[r] = me | 1u;
    if (i == 0)
        iparent[0] = k_invalid;
}

__device__ inline aabb aabb_union_ref(aabb const &a, aabb const &b) {
    return aabb{fmin_ref(a.min_x, b.min_x), fmin_ref(a.min_y, b.min_y), fmin_ref(a.min_z, b.min_z),
                fmax_ref(a.max_x, b.max_x), fmax_ref(a.max_y, b.max_y), fmax_ref(a.max_z, b.max_z)};
}

// Depth of every internal node of the Karras tree (edges to the root), by following the parent
// links: the parent array is small enough to stay in L2, and the tree's height bounds each walk.
// A walk longer than the node count means a broken parent chain: error.
__global__ __launch_bounds__(k_block) void lbvh_depth_kernel(
    std::uint32_t const *__restrict__ iparent, std::size_t ni, std::uint32_t *__restrict__ depth,
    std::uint32_t *__restrict__ max_depth, std::uint32_t *__restrict__ error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    std::uint32_t d = 0;
    if (i < ni) {
        std::uint32_t link = iparent[i];
        while (link != k_invalid) {
            if (++d > ni || (link >> 1) >= ni) {
                atomicOr(error, 1u);
                d = 0;
                break;
            }
            link = iparent[link >> 1];
        }
        depth[i] = d;
    }
    // Wave maximum first (every lane takes part), then one atomic per wave.
    for (int off = 32; off > 0; off >>= 1) {
        std::uint32_t const o = __shfl_xor(d, off);
        d = o > d ? o : d;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(max_depth, d);
}

// Bounds of the internal nodes at one depth, deepest depth first: both children are leaves or
// were written by an earlier launch, so each node unions them in left/right order (the
// reference's order) with no atomics or fences; the kernel boundary orders the depths.
__global__ __launch_bounds__(k_block) void lbvh_bounds_level_kernel(
    aabb const *__restrict__ leaf_boxes, uint2 const *__restrict__ bnodes, std::uint32_t const *__restrict__ depth,
    std::size_t ni, std::uint32_t level, aabb *__restrict__ ibounds, std::uint32_t *__restrict__ ifirst
) {
    std::size_t const node = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (node >= ni || depth[node] != level)
        return;
    uint2 const bn = bnodes[node];
    std::uint32_t const l = bn.x & ~k_leaf_bit, r = bn.y & ~k_leaf_bit;
    aabb const lb = (bn.x & k_leaf_bit) ? leaf_boxes[l] : ibounds[l];
    aabb const rb = (bn.y & k_leaf_bit) ? leaf_boxes[r] : ibounds[r];
    std::uint32_t const lf = (bn.x & k_leaf_bit) ? l : ifirst[l];
    std::uint32_t const rf = (bn.y & k_leaf_bit) ? r : ifirst[r];
    ibounds[node] = aabb_union_ref(lb, rb);
    ifirst[node] = lf < rf ? lf : rf;
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
// a < b creates internal node b - 1.  Each leaf is the larger-side minimum of at most one merge,
// so ids are unique, cover [0, T - 2], and do not depend on scheduling; bounds and first leaves
// are written beside the node, so no bounds climb follows.  Release before the boundary exchange
// and acquire after it publish a range's list and nodes to the lane (on any XCD) that joins it.
constexpr std::uint32_t k_hploc_list = 16; // clusters kept per partial range (hploc.cuh:23)
constexpr int k_hploc_block = 256;

__device__ inline std::uint64_t hploc_delta(std::uint64_t const *codes, std::uint32_t a, std::uint32_t b) {
    std::uint64_t const d = codes[a] ^ codes[b];
    return d == 0 ? static_cast<std::uint64_t>(a ^ b) : d;
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
    std::uint32_t *ifirst;
    std::uint32_t *clusters;
};

// PLOC over one range [L, R1) split at S, run by the 32 lanes of this half (all active, uniform
// control flow within the half).  Returns the merges made; sets `bad` on an inconsistent state.
__device__ inline std::uint32_t hploc_range(
    aabb const *__restrict__ leaf_boxes, std::uint32_t T, std::uint32_t radius, hploc_out const &o,
    std::uint32_t L, std::uint32_t R1, std::uint32_t S, bool fin, std::uint32_t *root_word, bool &bad
) {
    std::uint32_t const h = threadIdx.x & 31u;
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
        if (idx < T) {
            first = idx;
            box = leaf_boxes[idx];
        } else if (idx - T < T - 1u) {
            first = o.ifirst[idx - T];
            box = o.ibounds[idx - T];
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
            std::uint32_t const id = (first > n_first ? first : n_first) - 1u;
            if (first == n_first || id >= T - 1u) {
                bad = true;
            } else {
                aabb const merged = aabb_union_ref(box, n_box);
                std::uint32_t const f = first < n_first ? first : n_first;
                o.bnodes[id] = make_uint2(idx < T ? (idx | k_leaf_bit) : idx - T,
                                          n_idx < T ? (n_idx | k_leaf_bit) : n_idx - T);
                o.ibounds[id] = merged;
                o.ifirst[id] = f;
                idx = T + id;
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
        *root_word = idx - T;
    return made;
}

// words: [0] merges, [1] failure flag, [2] root internal id.
__global__ __launch_bounds__(k_hploc_block) void hploc_kernel(
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t T,
    std::uint32_t radius, hploc_out o, std::uint32_t *boundary, std::uint32_t *words
) {
    std::uint32_t const lane = threadIdx.x & 63u;
    std::uint32_t const idx = blockIdx.x * k_hploc_block + threadIdx.x;
    std::uint32_t left = idx, right = idx, split = 0;
    bool active = idx < T;
    bool bad = false;
    std::uint32_t made = 0;
    for (;;) {
        release_agent();
        if (active) {
            bool const to_right =
                left == 0u || (right != T - 1u && hploc_delta(codes, right, right + 1u) <
                                                        hploc_delta(codes, left - 1u, left));
            std::uint32_t const slot = to_right ? right : left - 1u;
            std::uint32_t const prev = __hip_atomic_exchange(
                &boundary[slot], to_right ? left : right, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        acquire_agent();
        bool const fin = active && right - left + 1u == T;
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
            if (sel >= 0)
                made += hploc_range(leaf_boxes, T, radius, o, rl, rr + 1u, rs, rf != 0, words + 2, bad);
        }
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

__global__ __launch_bounds__(k_block) void hploc_init_kernel(
    std::uint32_t *clusters, std::uint32_t *boundary, std::uint32_t T
) {
    std::uint32_t const i = blockIdx.x * k_block + threadIdx.x;
    if (i < T) {
        clusters[i] = i;
        boundary[i] = k_invalid;
    }
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

// Seeds level 0 with the binary root: node 0 for LBVH; for H-PLOC (hploc_words non-null) the
// recorded root, after checking that the build made T - 1 merges without failure.
__global__ void collapse_init_kernel(
    collapse_task *tasks, collapse_level *levels, collapse_stats *stats, std::uint32_t *parent,
    std::uint32_t const *hploc_words, std::uint32_t internal_count, aabb const *ibounds, aabb *root_box
) {
    std::uint32_t root = 0u, error = 0u;
    if (hploc_words != nullptr) {
        root = hploc_words[2];
        if (hploc_words[0] != internal_count || hploc_words[1] != 0u || root >= internal_count) {
            error = 1u;
            root = 0u;
        }
    }
    tasks[0] = collapse_task{root, 0u, 0u, 0u};
    // A failed H-PLOC leaves level 0 empty: the collapse launches then do nothing.
    levels[0] = collapse_level{0u, error ? 0u : 1u, 0u, 0u};
    *stats = collapse_stats{0u, 0u, 0u, error};
    parent[0] = k_invalid;
    *root_box = ibounds[root];
}

__device__ inline std::uint32_t wave_max_u32(std::uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        std::uint32_t const w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

__global__ __launch_bounds__(k_block) void collapse_level_kernel(
    uint2 const *__restrict__ bnodes, aabb const *__restrict__ ibounds,
    std::uint32_t const *__restrict__ ifirst, aabb const *__restrict__ leaf_boxes, int depth,
    std::uint32_t capacity, collapse_level *levels, collapse_task *tasks, std::uint32_t *status_cur,
    std::uint32_t *status_next, std::uint32_t status_words, collapse_stats *stats,
    bvh_node4 *__restrict__ nodes, std::uint32_t *__restrict__ parent,
    std::uint8_t *__restrict__ internal_children, int width
) {
    __shared__ std::uint32_t s_tile;
    __shared__ std::uint32_t s_exclusive;
    __shared__ std::uint32_t s_wave[k_block / 64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    // Only the first ceil(count / block) blocks take part (the launch is sized for the level's
    // bound, not its count): they claim exactly the tiles 0..tiles-1, each claim made by a running
    // block, so the look-back below never waits on an unclaimed tile.  A level past the last has
    // count 0 (the level table starts zeroed) and every block returns at once; no launch reads
    // the next status array then.
    std::uint64_t const level_tiles = (static_cast<std::uint64_t>(levels[depth].count) + k_block - 1) / k_block;
    if (blockIdx.x >= level_tiles)
        return;
    std::uint32_t const live = level_tiles < gridDim.x ? static_cast<std::uint32_t>(level_tiles) : gridDim.x;
    if (threadIdx.x == 0)
        s_tile = atomicAdd(&levels[depth].tiles_claimed, 1u);
    // The next level scans over the other status array; the live blocks clear it for that launch.
    for (std::uint32_t i = blockIdx.x * k_block + threadIdx.x; i < status_words; i += live * k_block)
        status_next[i] = 0u;
    __syncthreads();
    std::uint32_t const start = levels[depth].start;
    std::uint32_t const count = levels[depth].count;
    std::uint32_t const tile = s_tile;
    // A level never exceeds the launch (width^depth bound) or the capacity; report, do not index.
    if (static_cast<std::uint64_t>(start) + count > capacity ||
        count > static_cast<std::uint64_t>(gridDim.x) * k_block) {
        if (threadIdx.x == 0)
            atomicOr(&stats->error, 1u);
        return;
    }
    if (tile >= (count + k_block - 1) / k_block)
        return;
    std::uint32_t const i = tile * k_block + threadIdx.x;
    bool const active = i < count;
    std::uint32_t const self = start + i;

    std::uint32_t cand[4] = {k_invalid, k_invalid, k_invalid, k_invalid};
    int n_cand = 0;
    collapse_task task{};
    if (active) {
        task = tasks[self];
        uint2 const root = bnodes[task.binary];
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
                float const area = half_area(ibounds[cand[s]]);
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
            uint2 const ex = bnodes[cand[best]];
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
            first[s] = ifirst[cand[s]];
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
        std::uint32_t const ex = wave_lookback(status_cur, tile, block_total, &stats->error);
        if (lane == 0)
            s_exclusive = ex;
    }
    __syncthreads();
    std::uint32_t const next_start = start + count;
    if (threadIdx.x == 0 && tile == (count - 1) / k_block) {
        levels[depth + 1].start = next_start;
        levels[depth + 1].count = s_exclusive + block_total;
    }
    if (active) {
        std::uint32_t const base = next_start + s_exclusive + wave_prefix + incl - n_internal;
        std::uint32_t const node_internal = task.prefix_internal + (n_internal ? n_internal - 1u : 0u);
        std::uint32_t const node_packed = task.prefix_packed + static_cast<std::uint32_t>(n_cand - 1);
        bvh_node4 out{};
        for (int s = 0; s < n_cand; ++s) {
            std::uint32_t const c = cand[s] & ~k_leaf_bit;
            if (cand[s] & k_leaf_bit) {
                out.children[s].bounds = leaf_boxes[c];
                out.children[s].reference = encode_leaf(c, 1u);
                continue;
            }
            std::uint32_t rank = 0;
            for (int t = 0; t < 4; ++t)
                rank += first[t] < first[s] ? 1u : 0u;
            std::uint32_t const id = base + rank;
            if (id >= capacity) {
                atomicOr(&stats->error, 1u);
                continue;
            }
            out.children[s].bounds = ibounds[c];
            out.children[s].reference = encode_internal(id);
            tasks[id] = collapse_task{c, node_internal, node_packed, 0u};
            parent[id] = (self << 2) | static_cast<std::uint32_t>(s);
        }
        nodes[self] = out;
        internal_children[self] = static_cast<std::uint8_t>(n_internal);
        task.prefix_internal = node_internal;
        task.prefix_packed = node_packed;
    }
    std::uint32_t const w_int = wave_max_u32(active ? task.prefix_internal : 0u);
    std::uint32_t const w_pack = wave_max_u32(active ? task.prefix_packed : 0u);
    std::uint32_t const w_leaf = wave_sum_u32(n_leaf);
    if (lane == 0) {
        atomicMax(&stats->max_internal, w_int);
        atomicMax(&stats->max_packed, w_pack);
        atomicAdd(&stats->leaves, w_leaf);
    }
}

__global__ __launch_bounds__(k_block) void records_kernel(
    float const *__restrict__ vx, float const *__restrict__ vy, float const *__restrict__ vz,
    std::size_t nv, std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, std::uint32_t const *__restrict__ order, std::size_t nt,
    tri_record *__restrict__ records, std::uint32_t *__restrict__ error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= nt)
        return;
    std::uint32_t const p = order[i];
    std::uint32_t const a = i0[p], b = i1[p], c = i2[p];
    if (a >= nv || b >= nv || c >= nv) {
        atomicOr(error, 1u);
        return;
    }
    records[i] = tri_record{vx[a], vy[a], vz[a], vx[b], vy[b], vz[b], vx[c], vy[c], vz[c]};
}


} // namespace

void build_bvh4(gwn_geometry_handle const &g, gwn_bvh_handle &out, int method, std::uint32_t hploc_radius,
                hipStream_t stream, int width) {
    if (width < 2 || width > 4)
        throw_invalid("bvh.build: BVH width must be 2, 3 or 4.");
    std::size_t const T = g.triangle_count;
    std::size_t const V = g.vertex_count;
    out.triangle_count = T;
    out.node_count = 0;
    out.width = width;
    if (T == 0)
        return;

    // Scene bounds.
    device_buffer<std::uint32_t> scene_words(6 + 9, stream);
    float *scene = reinterpret_cast<float *>(scene_words.ptr + 6);
    {
        scene_init_kernel<<<1, 1, 0, stream>>>(scene_words.ptr);
        GWN_LAUNCH_CHECK();
        unsigned grid = grid_for(V, k_block);
        grid = std::min(grid, 512u);
        scene_bounds_kernel<<<grid, k_block, 0, stream>>>(g.x.ptr, g.y.ptr, g.z.ptr, V, scene_words.ptr);
        GWN_LAUNCH_CHECK();
        scene_finalize_kernel<<<1, 1, 0, stream>>>(scene_words.ptr, scene);
        GWN_LAUNCH_CHECK();
    }

    // AABB + Morton, stable sort, sorted boxes.
    device_buffer<aabb> boxes(T, stream);
    device_buffer<std::uint64_t> codes(T, stream), codes_alt(T, stream);
    device_buffer<std::uint32_t> ids(T, stream), ids_alt(T, stream);
    morton_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
        g.x.ptr, g.y.ptr, g.z.ptr, V, g.i0.ptr, g.i1.ptr, g.i2.ptr, T, scene, boxes.ptr, codes.ptr, ids.ptr
    );
    GWN_LAUNCH_CHECK();
    std::uint64_t *sorted_codes = nullptr;
    std::uint32_t *sorted_ids = nullptr;
    radix_sort_pairs_u64(codes.ptr, codes_alt.ptr, ids.ptr, ids_alt.ptr, T, 0, 63, &sorted_codes, &sorted_ids, stream);
    device_buffer<aabb> leaf_boxes(T, stream);
    gather_boxes_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(boxes.ptr, sorted_ids, T, leaf_boxes.ptr);
    GWN_LAUNCH_CHECK();

    // Primitive order and leaf-ordered records.
    out.primitive_indices.allocate(T, stream);
    GWN_HIP(hipMemcpyAsync(out.primitive_indices.ptr, sorted_ids, T * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
    out.triangles.allocate(T, stream);
    device_buffer<std::uint32_t> flags(1, stream);
    fill_u32(flags.ptr, 0u, 1, stream);
    records_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
        g.x.ptr, g.y.ptr, g.z.ptr, V, g.i0.ptr, g.i1.ptr, g.i2.ptr, out.primitive_indices.ptr, T,
        out.triangles.ptr, flags.ptr
    );
    GWN_LAUNCH_CHECK();

    if (T == 1) {
        aabb root_box{};
        std::uint32_t err = 0;
        GWN_HIP(hipMemcpyAsync(&root_box, leaf_boxes.ptr, sizeof(aabb), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipMemcpyAsync(&err, flags.ptr, sizeof(err), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        if (err != 0u)
            throw_internal("bvh.build: Leaf-ordered triangle construction encountered invalid geometry.");
        out.root.bounds = root_box;
        out.root.reference = encode_leaf(0, 1);
        out.internal_stack_bound = 0;
        out.packed_stack_bound = 0;
        out.max_depth = 0;
        return;
    }

    // Binary hierarchy with internal bounds and first leaves: Karras LBVH + bounds climb, or H-PLOC.
    std::size_t const NI = T - 1;
    device_buffer<uint2> bnodes(NI, stream);
    device_buffer<std::uint32_t> ifirst(NI, stream);
    device_buffer<aabb> ibounds(NI, stream);
    device_buffer<std::uint32_t> hploc_words;
    if (method == GWN_BVH_BUILD_HPLOC) {
        device_buffer<std::uint32_t> clusters(T, stream), boundary(T, stream);
        hploc_words.allocate(3, stream);
        fill_u32(hploc_words.ptr, 0u, 3, stream);
        // Every node id is written exactly once by a correct run; the fill keeps any node a
        // failed run leaves unwritten in bounds (two references to leaf 0) for the collapse.
        fill_u32(reinterpret_cast<std::uint32_t *>(bnodes.ptr), k_leaf_bit, 2 * NI, stream);
        fill_u32(ifirst.ptr, 0u, NI, stream);
        hploc_init_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
            clusters.ptr, boundary.ptr, static_cast<std::uint32_t>(T));
        GWN_LAUNCH_CHECK();
        hploc_kernel<<<grid_for(T, k_hploc_block), k_hploc_block, 0, stream>>>(
            sorted_codes, leaf_boxes.ptr, static_cast<std::uint32_t>(T), hploc_radius,
            hploc_out{bnodes.ptr, ibounds.ptr, ifirst.ptr, clusters.ptr}, boundary.ptr, hploc_words.ptr
        );
        GWN_LAUNCH_CHECK();
    } else {
        device_buffer<std::uint32_t> iparent(NI, stream), lparent(T, stream), depth(NI, stream), words(2, stream);
        fill_u32(words.ptr, 0u, 2, stream);
        lbvh_kernel<<<grid_for(NI, k_block), k_block, 0, stream>>>(sorted_codes, T, bnodes.ptr, iparent.ptr, lparent.ptr);
        GWN_LAUNCH_CHECK();
        // Bounds level by level, deepest first: the depths come from the parent links, and one
        // read-back of the tree's height sizes the loop (no leaf-to-root climb, no fences).
        lbvh_depth_kernel<<<grid_for(NI, k_block), k_block, 0, stream>>>(iparent.ptr, NI, depth.ptr, words.ptr,
                                                                         words.ptr + 1);
        GWN_LAUNCH_CHECK();
        std::uint32_t host_words[2] = {0, 0};
        GWN_HIP(hipMemcpyAsync(host_words, words.ptr, sizeof(host_words), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        if (host_words[1] != 0u)
            throw_internal("bvh.build: LBVH parent links do not form a tree.");
        for (std::uint32_t d = host_words[0] + 1; d-- > 0;) {
            lbvh_bounds_level_kernel<<<grid_for(NI, k_block), k_block, 0, stream>>>(
                leaf_boxes.ptr, bnodes.ptr, depth.ptr, NI, d, ibounds.ptr, ifirst.ptr);
            GWN_LAUNCH_CHECK();
        }
    }

    // Level-synchronous collapse into oversized arrays (wide nodes <= binary internal nodes).
    std::size_t const capacity = NI;
    std::uint32_t const status_words = static_cast<std::uint32_t>((capacity + k_block - 1) / k_block);
    device_buffer<collapse_task> tasks(capacity, stream);
    device_buffer<collapse_level> levels(k_max_levels, stream);
    device_buffer<collapse_stats> stats(1, stream);
    device_buffer<std::uint32_t> status(2 * static_cast<std::size_t>(status_words), stream);
    device_buffer<bvh_node4> wide(capacity, stream);
    device_buffer<std::uint32_t> wparent(capacity, stream);
    device_buffer<std::uint8_t> wchildren(capacity, stream);
    device_buffer<aabb> root_box_dev(1, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(levels.ptr), 0u, k_max_levels * 4, stream);
    fill_u32(status.ptr, 0u, 2 * static_cast<std::size_t>(status_words), stream);
    collapse_init_kernel<<<1, 1, 0, stream>>>(tasks.ptr, levels.ptr, stats.ptr, wparent.ptr, hploc_words.ptr,
                                              static_cast<std::uint32_t>(NI), ibounds.ptr, root_box_dev.ptr);
    GWN_LAUNCH_CHECK();

    // Launch levels in chunks and read the level table back after each chunk; a launch for a
    // level past the last one finds an empty level and returns.
    std::vector<collapse_level> host_levels(k_max_levels);
    int depth = 0;
    int chunk = 8;
    int empty_level = -1;
    while (empty_level < 0) {
        int const chunk_end = std::min(depth + chunk, k_max_levels - 1);
        for (; depth < chunk_end; ++depth) {
            // A level holds at most width^depth nodes and never more than the capacity.
            std::size_t bound = 1;
            for (int d = 0; d < depth && bound < capacity; ++d)
                bound *= static_cast<std::size_t>(width);
            bound = std::min<std::size_t>(bound, capacity);
            unsigned const grid = grid_for(bound, k_block);
            std::uint32_t *cur = status.ptr + static_cast<std::size_t>(depth & 1) * status_words;
            std::uint32_t *nxt = status.ptr + static_cast<std::size_t>((depth + 1) & 1) * status_words;
            collapse_level_kernel<<<grid, k_block, 0, stream>>>(
                bnodes.ptr, ibounds.ptr, ifirst.ptr, leaf_boxes.ptr, depth,
                static_cast<std::uint32_t>(capacity), levels.ptr, tasks.ptr, cur,
                nxt, status_words, stats.ptr, wide.ptr, wparent.ptr, wchildren.ptr, width
            );
            GWN_LAUNCH_CHECK();
        }
        GWN_HIP(hipMemcpyAsync(host_levels.data(), levels.ptr, static_cast<std::size_t>(depth + 1) * sizeof(collapse_level),
                               hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        for (int d = 0; d <= depth; ++d) {
            if (host_levels[d].count == 0u) {
                empty_level = d;
                break;
            }
        }
        if (empty_level < 0 && depth >= k_max_levels - 1)
            throw_internal("bvh.build.collapse: Canonical wide collapse exceeded the depth limit.");
        chunk *= 2;
    }
    collapse_stats hs{};
    aabb root_box{};
    std::uint32_t rec_err = 0;
    GWN_HIP(hipMemcpyAsync(&hs, stats.ptr, sizeof(hs), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&root_box, root_box_dev.ptr, sizeof(aabb), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&rec_err, flags.ptr, sizeof(rec_err), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (rec_err != 0u)
        throw_internal("bvh.build: Leaf-ordered triangle construction encountered invalid geometry.");
    std::size_t const NW = host_levels[empty_level].start;
    if (method == GWN_BVH_BUILD_HPLOC) {
        std::uint32_t hw[3] = {0, 0, 0};
        GWN_HIP(hipMemcpyAsync(hw, hploc_words.ptr, sizeof(hw), hipMemcpyDeviceToHost, stream));
        GWN_HIP(hipStreamSynchronize(stream));
        if (hw[0] != NI || hw[1] != 0u || hw[2] >= NI)
            throw_internal("bvh.build.hploc: H-PLOC builder did not converge to a complete binary hierarchy.");
    }
    if (hs.error != 0u || empty_level == 0 || NW == 0u || NW > capacity || hs.leaves != T)
        throw_internal("bvh.build.collapse: Canonical wide collapse reported an invalid output.");

    out.nodes.allocate(NW, stream);
    out.parent.allocate(NW, stream);
    out.internal_children.allocate(NW, stream);
    GWN_HIP(hipMemcpyAsync(out.nodes.ptr, wide.ptr, NW * sizeof(bvh_node4), hipMemcpyDeviceToDevice, stream));
    GWN_HIP(hipMemcpyAsync(out.parent.ptr, wparent.ptr, NW * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
    GWN_HIP(hipMemcpyAsync(out.internal_children.ptr, wchildren.ptr, NW, hipMemcpyDeviceToDevice, stream));

    out.node_count = NW;
    out.root.bounds = root_box;
    out.root.reference = encode_internal(0);
    out.internal_stack_bound = hs.max_internal;
    out.packed_stack_bound = hs.max_packed;
    out.max_depth = static_cast<std::uint32_t>(empty_level - 1);
    out.level_start.resize(static_cast<std::size_t>(empty_level) + 1);
    for (int d = 0; d <= empty_level; ++d)
        out.level_start[d] = host_levels[d].start;
}

} // namespace gwn
