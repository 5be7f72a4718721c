// gwn_build.hip - canonical BVH4 construction on gfx950.
//
// Pipeline (reference: include/gwn/detail/gwn_bvh_build_impl.cuh:44-255):
//   scene bounds (device reduction, no host read-back)       build_common.cuh:243-274
//   triangle AABB + 63-bit Morton code                       build_common.cuh:157-214
//   stable onesweep radix sort of (code, triangle id)        build_common.cuh:325-347
//   Karras LBVH                                              build_binary.cuh:29-129
//   bottom-up bounds + first leaf (arrival counters)         build_binary.cuh:131-192
//   BVH4 collapse on a persistent work queue                 build_collapse.cuh:98-314
//   breadth-first numbering by (depth, first sorted leaf)    (replaces reorder.cuh:61-143)
//   leaf-ordered triangle records                            build_impl.cuh:23-41
// One host synchronization remains, after the collapse, to size the node array from the
// device-side node count (the reference performs the same read-back, collapse.cuh:425-446).
#include "gwn_core.hpp"

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

// Per-axis min/max over all vertices; one atomic per wave per extremum.
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
    if ((threadIdx.x & 63) == 0) {
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

// Leaf-to-root climb: the second child to arrive at a node unions both children in left/right
// order, so the result is independent of arrival order.  Counters reset themselves to zero.
__global__ __launch_bounds__(k_block) void bounds_kernel(
    aabb const *__restrict__ leaf_boxes, uint2 const *__restrict__ bnodes,
    std::uint32_t const *__restrict__ iparent, std::uint32_t const *__restrict__ lparent,
    std::size_t n, aabb *ibounds, std::uint32_t *ifirst, std::uint32_t *arrivals
) {
    std::size_t const leaf = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (leaf >= n || n < 2)
        return;
    std::uint32_t link = lparent[leaf];
    while (link != k_invalid) {
        std::uint32_t const node = link >> 1;
        std::uint32_t const prior =
            __hip_atomic_fetch_add(&arrivals[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0u)
            return;
        arrivals[node] = 0u;
        uint2 const bn = bnodes[node];
        std::uint32_t const l = bn.x & ~k_leaf_bit, r = bn.y & ~k_leaf_bit;
        aabb const lb = (bn.x & k_leaf_bit) ? leaf_boxes[l] : ibounds[l];
        aabb const rb = (bn.y & k_leaf_bit) ? leaf_boxes[r] : ibounds[r];
        std::uint32_t const lf = (bn.x & k_leaf_bit) ? l : ifirst[l];
        std::uint32_t const rf = (bn.y & k_leaf_bit) ? r : ifirst[r];
        ibounds[node] = aabb_union_ref(lb, rb);
        ifirst[node] = lf < rf ? lf : rf;
        link = iparent[node];
    }
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

constexpr std::uint32_t k_collapse_idle_limit = 1u << 22;

struct collapse_task {
    std::uint32_t binary;
    std::uint32_t depth;
    std::uint32_t prefix_internal;
    std::uint32_t prefix_packed;
};

struct collapse_counters {
    std::uint32_t head;       // next queue position to claim
    std::uint32_t tail;       // next wide-node id to allocate (= queue length)
    std::uint32_t leaves_done;
    std::uint32_t max_internal;
    std::uint32_t max_packed;
    std::uint32_t max_depth;
    std::uint32_t error;
    std::uint32_t pad;
};

__global__ void collapse_init_kernel(collapse_task *tasks, std::uint32_t *ready, collapse_counters *ctr) {
    tasks[0] = collapse_task{0u, 0u, 0u, 0u};
    ready[0] = 1u;
    *ctr = collapse_counters{0u, 1u, 0u, 0u, 0u, 0u, 0u, 0u};
}

// Persistent BVH4 collapse.  Every wide node is one queue entry; its id is its queue position.
// A worker claims the next position, polls until the entry is published, expands the binary
// subtree to at most four candidates by always opening the internal candidate with the largest
// half-area (ties: smaller binary index; collapse.cuh:170-211), publishes internal children as
// new entries and retires.  There is no nested spin: a lane whose entry is not yet published
// falls through to the next loop iteration, so producer lanes of the same wave keep running.
// All wide nodes are done when every triangle has been emitted as a leaf.
__global__ __launch_bounds__(k_block) void collapse_kernel(
    uint2 const *__restrict__ bnodes, aabb const *__restrict__ ibounds,
    std::uint32_t const *__restrict__ ifirst, aabb const *__restrict__ leaf_boxes,
    std::uint32_t n_leaves, std::uint32_t capacity, int key_shift, collapse_task *tasks,
    std::uint32_t *ready, collapse_counters *ctr, bvh_node4 *__restrict__ wide,
    std::uint64_t *__restrict__ keys
) {
    std::uint32_t claim = k_invalid;
    std::uint32_t idle = 0;
    for (;;) {
        if (claim == k_invalid)
            claim = atomicAdd(&ctr->head, 1u);
        bool worked = false;
        if (claim < capacity &&
            __hip_atomic_load(&ready[claim], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
            worked = true;
            std::uint32_t const self = claim;
            claim = k_invalid;
            collapse_task const task = tasks[self];
            uint2 const root = bnodes[task.binary];
            std::uint32_t cand[4] = {root.x, root.y, k_invalid, k_invalid};
            int count = 2;
            while (count < 4) {
                int best = -1;
                float best_area = -INFINITY;
                std::uint32_t best_index = k_invalid;
                for (int s = 0; s < count; ++s) {
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
                cand[count] = ex.y;
                ++count;
            }
            std::uint32_t n_internal = 0, n_leaf = 0;
            for (int s = 0; s < count; ++s) {
                if (cand[s] & k_leaf_bit)
                    ++n_leaf;
                else
                    ++n_internal;
            }
            std::uint32_t const node_internal = task.prefix_internal + (n_internal ? n_internal - 1u : 0u);
            std::uint32_t const node_packed = task.prefix_packed + static_cast<std::uint32_t>(count - 1);
            bvh_node4 out{};
            for (int s = 0; s < count; ++s) {
                std::uint32_t const c = cand[s] & ~k_leaf_bit;
                if (cand[s] & k_leaf_bit) {
                    out.children[s].bounds = leaf_boxes[c];
                    out.children[s].reference = encode_leaf(c, 1u);
                    continue;
                }
                std::uint32_t const id = atomicAdd(&ctr->tail, 1u);
                if (id >= capacity) {
                    atomicOr(&ctr->error, 1u);
                    continue;
                }
                out.children[s].bounds = ibounds[c];
                out.children[s].reference = encode_internal(id);
                tasks[id] = collapse_task{c, task.depth + 1u, node_internal, node_packed};
                __hip_atomic_store(&ready[id], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            wide[self] = out;
            keys[self] = (static_cast<std::uint64_t>(task.depth) << key_shift) | ifirst[task.binary];
            atomicMax(&ctr->max_internal, node_internal);
            atomicMax(&ctr->max_packed, node_packed);
            atomicMax(&ctr->max_depth, task.depth);
            __hip_atomic_fetch_add(&ctr->leaves_done, n_leaf, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!worked) {
            std::uint32_t const done =
                __hip_atomic_load(&ctr->leaves_done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            std::uint32_t const err = __hip_atomic_load(&ctr->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (done >= n_leaves || err != 0u)
                break;
            // A well-formed binary tree always drains the queue; the idle cap only guarantees
            // that a malformed one ends the kernel (reported as an error) instead of hanging.
            if (++idle > k_collapse_idle_limit) {
                atomicOr(&ctr->error, 2u);
                break;
            }
        } else {
            idle = 0;
        }
        if (__ballot(worked) == 0ull)
            __builtin_amdgcn_s_sleep(2);
    }
}

__global__ __launch_bounds__(k_block) void make_keys32_kernel(
    std::uint64_t const *__restrict__ keys64, std::size_t n, std::uint32_t *__restrict__ keys32,
    std::uint32_t *__restrict__ ids
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < n) {
        keys32[i] = static_cast<std::uint32_t>(keys64[i]);
        ids[i] = static_cast<std::uint32_t>(i);
    }
}

__global__ __launch_bounds__(k_block) void iota_kernel(std::uint32_t *__restrict__ ids, std::size_t n) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < n)
        ids[i] = static_cast<std::uint32_t>(i);
}

__global__ __launch_bounds__(k_block) void invert_kernel(
    std::uint32_t const *__restrict__ perm, std::size_t n, std::uint32_t *__restrict__ inv
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < n)
        inv[perm[i]] = static_cast<std::uint32_t>(i);
}

// nodes[new] = wide[perm[new]] with internal child references renumbered; records parent links
// and internal-child counts for the bottom-up refits.
__global__ __launch_bounds__(k_block) void scatter_nodes_kernel(
    bvh_node4 const *__restrict__ wide, std::uint32_t const *__restrict__ perm,
    std::uint32_t const *__restrict__ inv, std::size_t n, bvh_node4 *__restrict__ nodes,
    std::uint32_t *__restrict__ parent, std::uint8_t *__restrict__ internal_children
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    bvh_node4 node = wide[perm[i]];
    std::uint8_t n_int = 0;
    for (int s = 0; s < 4; ++s) {
        std::uint64_t const r = node.children[s].reference;
        if (!ref_internal(r))
            continue;
        std::uint32_t const child = inv[ref_offset(r)];
        node.children[s].reference = encode_internal(child);
        parent[child] = (static_cast<std::uint32_t>(i) << 2) | static_cast<std::uint32_t>(s);
        ++n_int;
    }
    nodes[i] = node;
    internal_children[i] = n_int;
    if (i == 0)
        parent[0] = k_invalid;
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

int bits_for(std::uint64_t v) {
    int b = 0;
    while (b < 64 && (std::uint64_t(1) << b) < v)
        ++b;
    return b;
}

} // namespace

void build_bvh4(gwn_geometry_handle const &g, gwn_bvh_handle &out, hipStream_t stream) {
    std::size_t const T = g.triangle_count;
    std::size_t const V = g.vertex_count;
    out.triangle_count = T;
    out.node_count = 0;
    if (T == 0)
        return;

    // Scene bounds.
    device_buffer<std::uint32_t> scene_words(6 + 9, stream);
    float *scene = reinterpret_cast<float *>(scene_words.ptr + 6);
    {
        scene_init_kernel<<<1, 1, 0, stream>>>(scene_words.ptr);
        GWN_LAUNCH_CHECK();
        unsigned grid = grid_for(V, k_block);
        grid = std::min(grid, 1024u);
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

    // Karras LBVH + bounds.
    std::size_t const NI = T - 1;
    device_buffer<uint2> bnodes(NI, stream);
    device_buffer<std::uint32_t> iparent(NI, stream), lparent(T, stream), arrivals(NI, stream), ifirst(NI, stream);
    device_buffer<aabb> ibounds(NI, stream);
    fill_u32(arrivals.ptr, 0u, NI, stream);
    lbvh_kernel<<<grid_for(NI, k_block), k_block, 0, stream>>>(sorted_codes, T, bnodes.ptr, iparent.ptr, lparent.ptr);
    GWN_LAUNCH_CHECK();
    bounds_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
        leaf_boxes.ptr, bnodes.ptr, iparent.ptr, lparent.ptr, T, ibounds.ptr, ifirst.ptr, arrivals.ptr
    );
    GWN_LAUNCH_CHECK();

    // Persistent collapse.
    int const key_shift = bits_for(T);
    std::size_t const capacity = NI;
    device_buffer<collapse_task> tasks(capacity, stream);
    device_buffer<std::uint32_t> ready(capacity, stream);
    device_buffer<collapse_counters> ctr(1, stream);
    device_buffer<bvh_node4> wide(capacity, stream);
    device_buffer<std::uint64_t> keys(capacity, stream);
    fill_u32(ready.ptr, 0u, capacity, stream);
    collapse_init_kernel<<<1, 1, 0, stream>>>(tasks.ptr, ready.ptr, ctr.ptr);
    GWN_LAUNCH_CHECK();
    {
        int device = 0, cus = 0;
        GWN_HIP(hipGetDevice(&device));
        GWN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        std::size_t workers = static_cast<std::size_t>(std::max(cus, 1)) * 8u * k_block;
        workers = std::min(workers, ((capacity + k_block - 1) / k_block) * k_block);
        collapse_kernel<<<grid_for(workers, k_block), k_block, 0, stream>>>(
            bnodes.ptr, ibounds.ptr, ifirst.ptr, leaf_boxes.ptr, static_cast<std::uint32_t>(T),
            static_cast<std::uint32_t>(capacity), key_shift, tasks.ptr, ready.ptr, ctr.ptr, wide.ptr,
            keys.ptr
        );
        GWN_LAUNCH_CHECK();
    }
    collapse_counters hc{};
    aabb root_box{};
    std::uint32_t rec_err = 0;
    GWN_HIP(hipMemcpyAsync(&hc, ctr.ptr, sizeof(hc), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&root_box, ibounds.ptr, sizeof(aabb), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&rec_err, flags.ptr, sizeof(rec_err), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (rec_err != 0u)
        throw_internal("bvh.build: Leaf-ordered triangle construction encountered invalid geometry.");
    if (hc.error != 0u || hc.tail == 0u || hc.tail > capacity || hc.leaves_done != T)
        throw_internal("bvh.build.collapse: Canonical wide collapse reported an invalid output.");
    std::size_t const NW = hc.tail;

    // Breadth-first numbering: sort wide nodes by (depth, first sorted leaf).
    int const end_bit = key_shift + bits_for(static_cast<std::uint64_t>(hc.max_depth) + 1u);
    device_buffer<std::uint32_t> perm(NW, stream), perm_alt(NW, stream), inv(NW, stream);
    std::uint32_t *sorted_perm = nullptr;
    if (end_bit <= 32) {
        device_buffer<std::uint32_t> k32(NW, stream), k32_alt(NW, stream);
        make_keys32_kernel<<<grid_for(NW, k_block), k_block, 0, stream>>>(keys.ptr, NW, k32.ptr, perm.ptr);
        GWN_LAUNCH_CHECK();
        std::uint32_t *sk = nullptr;
        radix_sort_pairs_u32(k32.ptr, k32_alt.ptr, perm.ptr, perm_alt.ptr, NW, 0, end_bit, &sk, &sorted_perm, stream);
    } else {
        device_buffer<std::uint64_t> k64_alt(NW, stream);
        iota_kernel<<<grid_for(NW, k_block), k_block, 0, stream>>>(perm.ptr, NW);
        GWN_LAUNCH_CHECK();
        std::uint64_t *sk = nullptr;
        radix_sort_pairs_u64(keys.ptr, k64_alt.ptr, perm.ptr, perm_alt.ptr, NW, 0, end_bit, &sk, &sorted_perm, stream);
    }
    invert_kernel<<<grid_for(NW, k_block), k_block, 0, stream>>>(sorted_perm, NW, inv.ptr);
    GWN_LAUNCH_CHECK();
    out.nodes.allocate(NW, stream);
    out.parent.allocate(NW, stream);
    out.internal_children.allocate(NW, stream);
    scatter_nodes_kernel<<<grid_for(NW, k_block), k_block, 0, stream>>>(
        wide.ptr, sorted_perm, inv.ptr, NW, out.nodes.ptr, out.parent.ptr, out.internal_children.ptr
    );
    GWN_LAUNCH_CHECK();

    out.node_count = NW;
    out.root.bounds = root_box;
    out.root.reference = encode_internal(0);
    out.internal_stack_bound = hc.max_internal;
    out.packed_stack_bound = hc.max_packed;
    out.max_depth = hc.max_depth;
}

} // namespace gwn
