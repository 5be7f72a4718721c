// gwn_batch_build.hip - segmented builds: the BVH trees and Taylor moments of many meshes in one
// pass, for the batched mode's graph.
//
// No reference counterpart: the reference builds one mesh per call, with host read-backs inside
// (gwn_bvh_build_collapse.cuh:425-446, gwn_bvh_refit_moment_impl.cuh:260).  Here every mesh's
// triangles sit side by side in one set of arrays (mesh m owns sorted leaves [leaf_base, leaf_base
// + T) and binary / wide node slots [internal_base, internal_base + T - 1), internal_base =
// leaf_base - m), and one fixed sequence of launches builds all trees:
//   scene bounds per mesh            one launch over every mesh's vertex blocks
//   AABB + Morton code               one launch over all triangles
//   stable sort by (mesh, code)      the 63-bit code sort, then a stable sort by mesh id: within a
//                                    mesh, the order of a separate stable sort of that mesh alone
//   H-PLOC                           one launch over all sorted leaves; a mesh's first and last
//                                    leaves are hard range ends (hploc_climb)
//   BVH collapse, per wide level     a tile-prefix plan over the meshes, then one persistent
//                                    launch whose blocks claim tiles in order; each mesh's tiles
//                                    scan among themselves (collapse_tile)
//   moment refit, per wide level     segment tables planned on the device, then fixed-grid launches
//                                    (refit_moment_batch_device)
// Nothing is allocated and nothing is read back, so the pass can be captured into a hipGraph; the
// number of level launches is fixed by the caller and a deeper tree sets a status bit.  Each step
// runs the single-mesh build's device code (gwn_build_device.hpp) on the mesh's segment, so trees,
// records and moments equal those of a build and refit of each mesh alone, bit for bit.
#include "gwn_build_device.hpp"

#include <algorithm>
#include <vector>

namespace gwn {
namespace {

// One mesh of the segmented build (device table).
struct seg_mesh {
    float const *x, *y, *z;
    std::uint32_t const *i0, *i1, *i2;
    std::uint64_t vertex_count;
    std::uint32_t leaf_base, leaves, internal_base, pad;
};

// Last index j in [0, count) with prefix[j] <= v (prefix ascending, prefix[0] <= v).
template <class T>
__device__ inline std::uint32_t prefix_slot(T const *__restrict__ prefix, std::uint32_t count, T v) {
    std::uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        std::uint32_t const mid = lo + (hi - lo) / 2;
        if (prefix[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Exclusive scan of one value per thread over a 1024-thread block; *total gets the block's sum.
// Every thread calls it.
constexpr int k_plan_block = 1024;
__device__ inline std::uint64_t block_exclusive_scan(std::uint64_t v, std::uint64_t *total) {
    __shared__ std::uint64_t s_wave[k_plan_block / 64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    std::uint64_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        std::uint64_t const u = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += u;
    }
    if (lane == 63)
        s_wave[wave] = incl;
    __syncthreads();
    std::uint64_t before = 0, sum = 0;
    for (int w = 0; w < k_plan_block / 64; ++w) {
        if (w < wave)
            before += s_wave[w];
        sum += s_wave[w];
    }
    __syncthreads(); // s_wave is reused by the next call
    *total = sum;
    return before + incl - v;
}

__global__ __launch_bounds__(k_block) void seg_scene_init_kernel(std::uint32_t *ordered, std::uint32_t meshes) {
    std::uint32_t const m = blockIdx.x * k_block + threadIdx.x;
    if (m < meshes)
        scene_init_words(ordered + 6 * static_cast<std::size_t>(m));
}

// Block b belongs to the mesh whose vertex blocks [block_prefix[m], block_prefix[m + 1]) hold it.
__global__ __launch_bounds__(k_block) void seg_scene_bounds_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ block_prefix, std::uint32_t mesh_count,
    std::uint32_t *__restrict__ ordered
) {
    std::uint32_t const m = prefix_slot(block_prefix, mesh_count, static_cast<std::uint32_t>(blockIdx.x));
    seg_mesh const mm = meshes[m];
    std::uint32_t const local = blockIdx.x - block_prefix[m];
    std::uint32_t const blocks = block_prefix[m + 1] - block_prefix[m];
    scene_bounds_range(mm.x, mm.y, mm.z, static_cast<std::size_t>(local) * k_block + threadIdx.x, mm.vertex_count,
                       static_cast<std::size_t>(blocks) * k_block, ordered + 6 * static_cast<std::size_t>(m));
}

__global__ __launch_bounds__(k_block) void seg_scene_finalize_kernel(
    std::uint32_t const *ordered, float *scene, std::uint32_t meshes
) {
    std::uint32_t const m = blockIdx.x * k_block + threadIdx.x;
    if (m < meshes)
        scene_finalize_words(ordered + 6 * static_cast<std::size_t>(m), scene + 9 * static_cast<std::size_t>(m));
}

// Triangle g of the concatenation: its mesh's box and code; ids[g] = g.
__global__ __launch_bounds__(k_block) void seg_morton_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ leaf_prefix, std::uint32_t mesh_count,
    float const *__restrict__ scene, std::uint32_t total, aabb *__restrict__ boxes,
    std::uint64_t *__restrict__ codes, std::uint32_t *__restrict__ ids
) {
    std::uint32_t const g = blockIdx.x * k_block + threadIdx.x;
    if (g >= total)
        return;
    std::uint32_t const m = prefix_slot(leaf_prefix, mesh_count, g);
    seg_mesh const mm = meshes[m];
    aabb box;
    std::uint64_t code;
    morton_triangle(mm.x, mm.y, mm.z, mm.vertex_count, mm.i0, mm.i1, mm.i2, g - mm.leaf_base,
                    scene + 9 * static_cast<std::size_t>(m), box, code);
    boxes[g] = box;
    codes[g] = code;
    ids[g] = g;
}

// After the code sort: each position's mesh (the key of the stable sort by mesh) and the position.
__global__ __launch_bounds__(k_block) void seg_mesh_key_kernel(
    std::uint32_t const *__restrict__ sorted_ids, std::uint32_t const *__restrict__ leaf_prefix,
    std::uint32_t mesh_count, std::uint32_t total, std::uint32_t *__restrict__ keys, std::uint32_t *__restrict__ pos
) {
    std::uint32_t const p = blockIdx.x * k_block + threadIdx.x;
    if (p >= total)
        return;
    keys[p] = prefix_slot(leaf_prefix, mesh_count, sorted_ids[p]);
    pos[p] = p;
}

// Sorted position i (mesh-major, code order within the mesh): its code, leaf box, the mesh-local
// primitive index and the leaf-ordered triangle record.  `mesh_of` / `pos_of` are the mesh sort's
// output (null with one mesh: the code sort's order is final).  error: 1 bad vertex index,
// 2 a position outside its mesh's range.
__global__ __launch_bounds__(k_block) void seg_gather_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t total, std::uint32_t const *__restrict__ mesh_of,
    std::uint32_t const *__restrict__ pos_of, std::uint64_t const *__restrict__ codes1,
    std::uint32_t const *__restrict__ ids1, aabb const *__restrict__ boxes, std::uint64_t *__restrict__ codes,
    aabb *__restrict__ leaf_boxes, std::uint32_t *__restrict__ prim, tri_record *__restrict__ tris,
    std::uint32_t *__restrict__ error
) {
    std::uint32_t const i = blockIdx.x * k_block + threadIdx.x;
    if (i >= total)
        return;
    std::uint32_t const p = pos_of ? pos_of[i] : i;
    std::uint32_t const m = mesh_of ? mesh_of[i] : 0u;
    std::uint32_t const g = ids1[p];
    seg_mesh const mm = meshes[m];
    codes[i] = codes1[p];
    leaf_boxes[i] = boxes[g];
    std::uint32_t const local = g - mm.leaf_base;
    prim[i] = local;
    if (i - mm.leaf_base >= mm.leaves || local >= mm.leaves) {
        atomicOr(error, 2u);
        return;
    }
    if (!triangle_record(mm.x, mm.y, mm.z, mm.vertex_count, mm.i0, mm.i1, mm.i2, local, tris + i))
        atomicOr(error, 1u);
}

__global__ __launch_bounds__(k_block) void seg_hploc_init_kernel(
    std::uint32_t *clusters, std::uint32_t *boundary, std::uint32_t total
) {
    std::uint32_t const i = blockIdx.x * k_block + threadIdx.x;
    if (i < total) {
        clusters[i] = i;
        boundary[i] = k_invalid;
    }
}

__device__ inline build_segment seg_of(seg_mesh const *__restrict__ meshes,
                                      std::uint32_t const *__restrict__ leaf_prefix, std::uint32_t mesh_count,
                                      std::uint32_t leaf) {
    std::uint32_t const m = prefix_slot(leaf_prefix, mesh_count, leaf);
    seg_mesh const mm = meshes[m];
    return build_segment{mm.leaf_base, mm.leaves, mm.internal_base, m};
}

// H-PLOC chunk launch: one lane per sorted leaf of every mesh (meshes of one triangle take no
// part); a chunk may hold several meshes, whose ends stay hard range ends.  words: [0] merges,
// [1] failure flag, [4] parked ranges.
__global__ __launch_bounds__(k_hploc_seg_chunk_block) void seg_hploc_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ leaf_prefix, std::uint32_t mesh_count,
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t total,
    std::uint32_t radius, hploc_out o, std::uint32_t *boundary, uint2 *parked, std::uint32_t *root_words,
    std::uint32_t *words
) {
    std::uint32_t const idx = blockIdx.x * k_hploc_seg_chunk_block + threadIdx.x;
    build_segment seg{0u, 0u, 0u, 0u};
    bool active = idx < total;
    if (active) {
        seg = seg_of(meshes, leaf_prefix, mesh_count, idx);
        active = seg.leaves >= 2u;
    }
    hploc_chunk_climb<k_hploc_seg_chunk_block>(codes, leaf_boxes, total, radius, o, boundary, seg, idx, active,
                                               parked, words + 4, root_words, words);
}

// H-PLOC global launch over the ranges the chunk launch parked.
__global__ __launch_bounds__(k_hploc_block) void seg_hploc_parked_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ leaf_prefix, std::uint32_t mesh_count,
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t total,
    std::uint32_t radius, hploc_out o, std::uint32_t *boundary, uint2 const *parked, std::uint32_t *root_words,
    std::uint32_t *words
) {
    std::uint32_t const i = blockIdx.x * k_hploc_block + threadIdx.x;  // one range per lane (they are many and small)
    bool const active = i < words[4];
    uint2 const r = active ? parked[i] : make_uint2(0u, 0u);
    build_segment const seg = active ? seg_of(meshes, leaf_prefix, mesh_count, r.x) : build_segment{0u, 0u, 0u, 0u};
    hploc_climb<false>(codes, leaf_boxes, total, radius, o, boundary, seg, r.x, r.y, active, 0u, 0u, nullptr, nullptr,
                       root_words, words);
}

// Karras LBVH over every mesh: one thread per binary internal node of the concatenation (mesh m's
// nodes are [internal_prefix[m], internal_prefix[m + 1]); meshes of one triangle have none).
__global__ __launch_bounds__(k_block) void seg_lbvh_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ internal_prefix, std::uint32_t mesh_count,
    std::uint32_t internal_total, std::uint64_t const *__restrict__ codes, uint2 *__restrict__ bnodes,
    std::uint32_t *__restrict__ iparent
) {
    std::uint32_t const j = blockIdx.x * k_block + threadIdx.x;
    if (j >= internal_total)
        return;
    seg_mesh const mm = meshes[prefix_slot(internal_prefix, mesh_count, j)];
    lbvh_node(codes + mm.leaf_base, mm.leaves, j - mm.internal_base, bnodes + mm.internal_base,
              iparent + mm.internal_base, nullptr);
}

// Binary depths of every mesh's internal nodes by pointer jumping.  jump[j] packs the ancestor
// (low word) and the edges to it (high word) of internal node j; a root is its own ancestor at
// distance 0.  Each round replaces (a, d) by (ancestor of a,
// d + a's distance), read as one 64-bit word, so every word stays a true (ancestor, distance)
// pair whatever order the lanes run in, and after r rounds a node has climbed at least
// min(depth, 2^r) edges.  Seven rounds reach 128 = k_batch_build_max_levels.
constexpr int k_jump_rounds = 7;
static_assert((1 << k_jump_rounds) >= k_batch_build_max_levels, "jump rounds must cover the level limit");

__device__ inline std::uint64_t jump_pack(std::uint32_t ancestor, std::uint32_t distance) {
    return static_cast<std::uint64_t>(ancestor) | (static_cast<std::uint64_t>(distance) << 32);
}

__global__ __launch_bounds__(k_block) void seg_jump_init_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ internal_prefix, std::uint32_t mesh_count,
    std::uint32_t internal_total, std::uint32_t const *__restrict__ iparent, std::uint64_t *__restrict__ jump,
    std::uint32_t *error
) {
    std::uint32_t const j = blockIdx.x * k_block + threadIdx.x;
    if (j >= internal_total)
        return;
    seg_mesh const mm = meshes[prefix_slot(internal_prefix, mesh_count, j)];
    std::uint32_t const link = iparent[j];
    if (link == k_invalid) {
        jump[j] = jump_pack(j, 0u);
    } else if ((link >> 1) >= mm.leaves - 1u) {
        atomicOr(error, 1u); // a parent outside the mesh: a broken tree
        jump[j] = jump_pack(j, 0u);
    } else {
        jump[j] = jump_pack(mm.internal_base + (link >> 1), 1u);
    }
}

__global__ __launch_bounds__(k_block) void seg_jump_kernel(std::uint32_t internal_total, std::uint64_t *jump) {
    std::uint32_t const j = blockIdx.x * k_block + threadIdx.x;
    if (j >= internal_total)
        return;
    std::uint64_t const p = __hip_atomic_load(jump + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    std::uint32_t const a = static_cast<std::uint32_t>(p);
    if (a == j)
        return;
    std::uint64_t const q = __hip_atomic_load(jump + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(jump + j, jump_pack(static_cast<std::uint32_t>(q), static_cast<std::uint32_t>(p >> 32) +
                                                                         static_cast<std::uint32_t>(q >> 32)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The contiguous share [lo, hi) of `total` items that block blockIdx.x of a fixed grid takes.
__device__ inline void block_share(std::uint32_t total, std::uint32_t &lo, std::uint32_t &hi) {
    std::uint64_t const per = (static_cast<std::uint64_t>(total) + gridDim.x - 1) / gridDim.x;
    std::uint64_t const a = per * blockIdx.x;
    lo = static_cast<std::uint32_t>(a < total ? a : total);
    hi = static_cast<std::uint32_t>(a + per < total ? a + per : total);
}

__device__ inline std::uint32_t depth_bin(std::uint32_t d) {
    return d < k_batch_build_max_levels ? d : k_batch_build_max_levels - 1;
}

// Depth of every node after the rounds, the deepest into *max_depth (all ones if some node did
// not reach its root: a tree deeper than the level limit), and the depth histogram into
// counts[depth] (depths clamp to the last bin).  A fixed grid: each block counts its share in LDS
// and adds it to the global counts once.
__global__ __launch_bounds__(k_block) void seg_depth_kernel(
    std::uint32_t internal_total, std::uint64_t const *__restrict__ jump, std::uint32_t *__restrict__ depth,
    std::uint32_t *counts, std::uint32_t *max_depth
) {
    __shared__ std::uint32_t s_hist[k_batch_build_max_levels];
    __shared__ std::uint32_t s_max;
    for (int b = threadIdx.x; b < k_batch_build_max_levels; b += k_block)
        s_hist[b] = 0u;
    if (threadIdx.x == 0)
        s_max = 0u;
    __syncthreads();
    std::uint32_t lo, hi;
    block_share(internal_total, lo, hi);
    std::uint32_t dmax = 0;
    for (std::uint32_t j = lo + threadIdx.x; j < hi; j += k_block) {
        std::uint64_t const p = jump[j];
        std::uint32_t const a = static_cast<std::uint32_t>(p);
        std::uint32_t d = static_cast<std::uint32_t>(p >> 32);
        if (static_cast<std::uint32_t>(jump[a]) != a)
            d = 0xffffffffu;
        depth[j] = d;
        dmax = d > dmax ? d : dmax;
        atomicAdd(&s_hist[depth_bin(d)], 1u);
    }
    for (int off = 32; off > 0; off >>= 1) {
        std::uint32_t const o = __shfl_xor(dmax, off);
        dmax = o > dmax ? o : dmax;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(&s_max, dmax);
    __syncthreads();
    for (int b = threadIdx.x; b < k_batch_build_max_levels; b += k_block)
        if (s_hist[b] != 0u)
            atomicAdd(&counts[b], s_hist[b]);
    if (threadIdx.x == 0 && hi > lo)
        atomicMax(max_depth, s_max);
}

// Bucket starts (exclusive prefix of counts, bucket_start[k_batch_build_max_levels] = total) and
// the scatter cursors cleared.  One block.
__global__ __launch_bounds__(k_batch_build_max_levels) void seg_depth_scan_kernel(
    std::uint32_t const *__restrict__ counts, std::uint32_t *__restrict__ bucket_start, std::uint32_t *cursor
) {
    __shared__ std::uint32_t s_c[k_batch_build_max_levels];
    std::uint32_t const b = threadIdx.x;
    s_c[b] = counts[b];
    __syncthreads();
    if (b == 0) {
        std::uint32_t run = 0;
        for (int k = 0; k < k_batch_build_max_levels; ++k) {
            bucket_start[k] = run;
            run += s_c[k];
        }
        bucket_start[k_batch_build_max_levels] = run;
    }
    cursor[b] = 0u;
}

// Node ids into their depth's bucket (any order within a bucket: each node's bounds are computed
// on their own).  A fixed grid over the same shares: each block counts its share per depth,
// reserves those places with one atomic per depth, then writes its nodes into them.
__global__ __launch_bounds__(k_block) void seg_depth_scatter_kernel(
    std::uint32_t internal_total, std::uint32_t const *__restrict__ depth, std::uint32_t const *__restrict__ bucket_start,
    std::uint32_t *cursor, std::uint32_t *__restrict__ order
) {
    __shared__ std::uint32_t s_count[k_batch_build_max_levels];
    __shared__ std::uint32_t s_base[k_batch_build_max_levels];
    for (int b = threadIdx.x; b < k_batch_build_max_levels; b += k_block)
        s_count[b] = 0u;
    __syncthreads();
    std::uint32_t lo, hi;
    block_share(internal_total, lo, hi);
    for (std::uint32_t j = lo + threadIdx.x; j < hi; j += k_block)
        atomicAdd(&s_count[depth_bin(depth[j])], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < k_batch_build_max_levels; b += k_block) {
        std::uint32_t const c = s_count[b];
        s_base[b] = bucket_start[b] + (c != 0u ? atomicAdd(&cursor[b], c) : 0u);
        s_count[b] = 0u;
    }
    __syncthreads();
    for (std::uint32_t j = lo + threadIdx.x; j < hi; j += k_block) {
        std::uint32_t const bin = depth_bin(depth[j]);
        order[s_base[bin] + atomicAdd(&s_count[bin], 1u)] = j;
    }
}

// Bounds of every mesh's internal nodes at binary depth `level` (launched deepest first): a fixed
// grid strides over that depth's bucket.
__global__ __launch_bounds__(k_block) void seg_lbvh_bounds_level_kernel(
    seg_mesh const *__restrict__ meshes, std::uint32_t const *__restrict__ internal_prefix, std::uint32_t mesh_count,
    std::uint32_t const *__restrict__ order, std::uint32_t const *__restrict__ bucket_start, std::uint32_t level,
    aabb const *__restrict__ leaf_boxes, uint2 const *__restrict__ bnodes, aabb *__restrict__ ibounds,
    std::uint32_t *__restrict__ ifirst
) {
    std::uint32_t const end = bucket_start[level + 1];
    for (std::uint32_t k = bucket_start[level] + blockIdx.x * k_block + threadIdx.x; k < end;
         k += gridDim.x * k_block) {
        std::uint32_t const j = order[k];
        seg_mesh const mm = meshes[prefix_slot(internal_prefix, mesh_count, j)];
        lbvh_node_bounds(leaf_boxes + mm.leaf_base, bnodes + mm.internal_base, j - mm.internal_base,
                         ibounds + mm.internal_base, ifirst + mm.internal_base);
    }
}

// Level 0 of every mesh's collapse: the binary root (LBVH: node 0; H-PLOC: the recorded root) of a
// tree of two or more leaves, or an empty level (one leaf: the root is that leaf).
__global__ __launch_bounds__(k_block) void seg_collapse_init_kernel(
    seg_mesh const *__restrict__ meshes, collapse_tree const *__restrict__ trees, std::uint32_t mesh_count,
    std::uint32_t const *__restrict__ root_words, aabb *__restrict__ root_box
) {
    std::uint32_t const m = blockIdx.x * k_block + threadIdx.x;
    if (m >= mesh_count)
        return;
    seg_mesh const mm = meshes[m];
    collapse_tree const t = trees[m];
    if (mm.leaves < 2u) {
        t.levels[0] = collapse_level{0u, 0u, 0u, 0u};
        *t.stats = collapse_stats{0u, 0u, mm.leaves, 0u};
        root_box[m] = t.leaf_boxes[0];
        return;
    }
    std::uint32_t root = root_words[m];
    std::uint32_t error = 0u;
    if (root >= mm.leaves - 1u) {
        error = 1u;
        root = 0u;
    }
    t.tasks[0] = collapse_task{root, 0u, 0u, 0u};
    // A failed H-PLOC leaves level 0 empty: the collapse launches then do nothing for this mesh.
    t.levels[0] = collapse_level{0u, error ? 0u : 1u, 0u, 0u};
    *t.stats = collapse_stats{0u, 0u, 0u, error};
    t.parent[0] = k_invalid;
    root_box[m] = t.ibounds[root];
}

// Node count of a level as far as the mesh's node slots hold it: a level table past the slots
// (only a broken binary tree can give one) counts as empty here, and the level launches report it.
__device__ inline std::uint32_t level_nodes(collapse_level const &lv, std::uint32_t capacity) {
    return static_cast<std::uint64_t>(lv.start) + lv.count <= capacity ? lv.count : 0u;
}

// Tiles of level `depth` per mesh, as an exclusive prefix (tile_prefix[mesh_count] = total, at most
// the status array's size), and the level's tile claim counter reset.
__global__ __launch_bounds__(k_plan_block) void seg_collapse_plan_kernel(
    collapse_tree const *__restrict__ trees, std::uint32_t mesh_count, int depth,
    std::uint32_t *__restrict__ tile_prefix, std::uint32_t *__restrict__ claim
) {
    std::uint64_t base = 0;
    for (std::uint32_t c = 0; c < mesh_count; c += k_plan_block) {
        std::uint32_t const m = c + threadIdx.x;
        std::uint64_t tiles = 0;
        if (m < mesh_count) {
            collapse_tree const &t = trees[m];
            tiles = (level_nodes(t.levels[depth], t.capacity) + k_block - 1) / k_block;
        }
        std::uint64_t sum = 0;
        std::uint64_t const ex = block_exclusive_scan(tiles, &sum);
        if (m < mesh_count)
            tile_prefix[m] = static_cast<std::uint32_t>(base + ex);
        base += sum;
    }
    if (threadIdx.x == 0) {
        tile_prefix[mesh_count] = static_cast<std::uint32_t>(base);
        *claim = 0u;
    }
}

// One wide level of every mesh.  Blocks claim the level's tiles in order from one counter (mesh
// by mesh, each mesh's tiles in order), so a tile's look-back only waits on tiles of its mesh that
// running blocks hold; each mesh scans within its own span of the status array.
__global__ __launch_bounds__(k_block) void seg_collapse_level_kernel(
    collapse_tree const *__restrict__ trees, std::uint32_t mesh_count, int depth,
    std::uint32_t const *__restrict__ tile_prefix, std::uint32_t *claim, std::uint32_t *status_cur,
    std::uint32_t *status_next, std::uint32_t status_words, int width
) {
    __shared__ std::uint32_t s_tile;
    std::uint32_t const total = tile_prefix[mesh_count];
    if (total == 0u)
        return; // every mesh is finished: so are all deeper levels
    // The next level scans over the other status array: clear it for that launch.
    for (std::uint32_t i = blockIdx.x * k_block + threadIdx.x; i < status_words; i += gridDim.x * k_block)
        status_next[i] = 0u;
    for (;;) {
        if (threadIdx.x == 0)
            s_tile = atomicAdd(claim, 1u);
        __syncthreads();
        std::uint32_t const g = s_tile;
        __syncthreads(); // every thread has read s_tile before the next claim writes it
        if (g >= total)
            return;
        std::uint32_t const m = prefix_slot(tile_prefix, mesh_count, g);
        collapse_tree const t = trees[m];
        std::uint32_t const start = t.levels[depth].start;
        std::uint32_t const count = t.levels[depth].count;
        // The plan gave this mesh tiles only if the level fits its node slots.
        collapse_tile(t, depth, start, count, g - tile_prefix[m], status_cur + tile_prefix[m], width);
    }
}

// Per mesh after the last level launch: a failed build (2), a tree deeper than the launched levels
// (4), a traversal stack bound above the compiled capacity (8).  Mesh 0's thread also checks the
// words: [0] H-PLOC merges, [1] failure flag (H-PLOC or an LBVH parent chain), [2] the records'
// error word, [3] the deepest LBVH binary level against the bounds launches.
__global__ __launch_bounds__(k_block) void seg_check_kernel(
    seg_mesh const *__restrict__ meshes, collapse_tree const *__restrict__ trees, std::uint32_t mesh_count,
    int levels, std::uint32_t stack_capacity, std::uint32_t const *__restrict__ words, std::uint32_t expected_merges,
    std::uint32_t binary_levels, std::uint32_t *status
) {
    std::uint32_t const m = blockIdx.x * k_block + threadIdx.x;
    if (m == 0u && (words[0] != expected_merges || words[1] != 0u || words[2] != 0u))
        atomicOr(status, 2u);
    if (m == 0u && words[3] >= binary_levels)
        atomicOr(status, 4u); // an LBVH deeper than the bounds launches
    if (m >= mesh_count)
        return;
    seg_mesh const mm = meshes[m];
    if (mm.leaves < 2u)
        return;
    collapse_tree const t = trees[m];
    collapse_stats const st = *t.stats;
    std::uint32_t bits = 0u;
    if (st.error != 0u || st.leaves != mm.leaves)
        bits |= 2u;
    for (int d = 0; d <= levels; ++d)
        if (static_cast<std::uint64_t>(t.levels[d].start) + t.levels[d].count > t.capacity)
            bits |= 2u; // a level past the node slots (the collapse skipped it)
    if (t.levels[levels].count != 0u)
        bits |= 4u;
    if (st.max_internal > stack_capacity)
        bits |= 8u;
    if (bits)
        atomicOr(status, bits);
}

// The moment refit's tables from the collapse's level tables: block d < levels lists the meshes
// with nodes at depth d (segments d * mesh_count ...) and their node count; block `levels` writes
// the slot prefix (4 x wide nodes per mesh) of the evaluation pass.
__global__ __launch_bounds__(k_plan_block) void seg_refit_plan_kernel(
    seg_mesh const *__restrict__ meshes, collapse_tree const *__restrict__ trees, std::uint32_t mesh_count,
    int levels, refit_segment *__restrict__ segments, refit_level *__restrict__ refit_levels,
    std::uint64_t *__restrict__ slot_prefix
) {
    int const d = static_cast<int>(blockIdx.x);
    std::uint64_t base_nodes = 0, base_segments = 0;
    for (std::uint32_t c = 0; c < mesh_count; c += k_plan_block) {
        std::uint32_t const m = c + threadIdx.x;
        bool const tree = m < mesh_count && meshes[m].leaves >= 2u;
        collapse_level const *lv = tree ? trees[m].levels : nullptr;
        std::uint32_t const capacity = tree ? trees[m].capacity : 0u;
        std::uint64_t nodes = 0;
        if (d < levels) {
            if (tree)
                nodes = level_nodes(lv[d], capacity);
            std::uint64_t n_sum = 0, s_sum = 0;
            std::uint64_t const n_ex = block_exclusive_scan(nodes, &n_sum);
            std::uint64_t const s_ex = block_exclusive_scan(nodes ? 1u : 0u, &s_sum);
            if (nodes)
                segments[static_cast<std::size_t>(d) * mesh_count + base_segments + s_ex] =
                    refit_segment{m, lv[d].start, base_nodes + n_ex};
            base_nodes += n_sum;
            base_segments += s_sum;
        } else {
            // Wide nodes of the mesh: the furthest end of any level (the first empty level starts
            // there).
            if (tree)
                for (int k = 0; k <= levels; ++k)
                    if (level_nodes(lv[k], capacity) != 0u)
                        nodes = std::max<std::uint64_t>(nodes, static_cast<std::uint64_t>(lv[k].start) + lv[k].count);
            std::uint64_t sum = 0;
            std::uint64_t const ex = block_exclusive_scan(4 * nodes, &sum);
            if (m < mesh_count)
                slot_prefix[m] = base_nodes + ex;
            base_nodes += sum;
        }
    }
    if (threadIdx.x == 0) {
        if (d < levels)
            refit_levels[d] = refit_level{base_nodes, static_cast<std::uint32_t>(d) * mesh_count,
                                          static_cast<std::uint32_t>(base_segments)};
        else
            slot_prefix[mesh_count] = base_nodes;
    }
}

} // namespace

struct batch_trees {
    std::size_t mesh_count = 0;
    std::uint32_t total = 0;           // triangles of all meshes
    std::uint32_t internal_total = 0;  // binary internal nodes of all meshes (= node slots)
    std::uint32_t vertex_blocks = 0;
    std::uint32_t status_words = 0;
    int method = GWN_BVH_BUILD_HPLOC;
    int width = 4;
    std::uint32_t radius = 8;
    int order = 1;
    int mesh_bits = 0;
    unsigned grid = 1;
    std::vector<seg_mesh> host_meshes;
    device_buffer<seg_mesh> meshes;
    device_buffer<std::uint32_t> leaf_prefix, block_prefix, internal_prefix;
    // Build scratch.
    device_buffer<std::uint32_t> scene_words;
    device_buffer<float> scene;
    device_buffer<aabb> boxes, leaf_boxes;
    device_buffer<std::uint64_t> codes, codes_alt;
    device_buffer<std::uint32_t> ids, ids_alt, keys, keys_alt, pos, pos_alt, sort_ws;
    device_buffer<uint2> bnodes;
    device_buffer<aabb> ibounds;
    device_buffer<std::uint32_t> ifirst, clusters, boundary, root_words, words, iparent, depth;
    device_buffer<uint2> parked;
    device_buffer<std::uint64_t> jump;
    device_buffer<std::uint32_t> by_depth, buckets;  // buckets: counts | starts (+1) | cursors
    device_buffer<collapse_task> tasks;
    device_buffer<collapse_level> levels;  // mesh_count x (k_batch_build_max_levels + 1)
    device_buffer<collapse_stats> stats;
    device_buffer<std::uint32_t> status, tile_prefix, claim;
    device_buffer<collapse_tree> trees;
    // The trees.
    device_buffer<bvh_node4> nodes;
    device_buffer<std::uint32_t> parent, prim;
    device_buffer<std::uint8_t> internal_children;
    device_buffer<tri_record> tris;
    device_buffer<aabb> root_box;
    // The moments and their refit tables.
    device_buffer<float4> records, eval_records;
    device_buffer<float> slot_area;
    device_buffer<std::uint8_t> pending;
    device_buffer<moment_refit_mesh> refit_meshes;
    device_buffer<refit_segment> segments;
    device_buffer<refit_level> refit_levels;
    device_buffer<std::uint64_t> slot_prefix;
    std::size_t bytes = 0;
};

void batch_trees_deleter::operator()(batch_trees *t) const noexcept { delete t; }

namespace {

constexpr std::uint32_t k_level_stride = k_batch_build_max_levels + 1;

template <class T> void alloc(device_buffer<T> &b, std::size_t n, hipStream_t stream, std::size_t &bytes) {
    b.allocate(n, stream);
    bytes += n * sizeof(T);
}

unsigned blocks_for(std::size_t n, int block) { return grid_for(n, block); }

} // namespace

batch_trees_ptr batch_trees_create(std::vector<gwn_geometry_handle const *> const &geometries, int method,
                                   int width, std::uint32_t hploc_radius, int order, hipStream_t stream) {
    if (order < 0 || order > 2)
        throw_invalid("Taylor order must be 0, 1, or 2.");
    batch_trees_ptr tp(new batch_trees());
    batch_trees &t = *tp;
    t.order = order;
    batch_trees_set_build(t, method, width, hploc_radius);
    std::size_t const M = geometries.size();
    t.mesh_count = M;
    if (M == 0)
        return tp;
    if (M > (1u << 24))
        throw_invalid("batch: Too many meshes for one segmented build.");
    std::uint64_t leaves = 0, blocks = 0, status_words = 0;
    std::vector<std::uint32_t> leaf_prefix(M + 1), block_prefix(M + 1), internal_prefix(M + 1);
    for (std::size_t m = 0; m < M; ++m) {
        gwn_geometry_handle const &g = *geometries[m];
        std::uint64_t const T = g.triangle_count;
        if (T == 0)
            throw_invalid("batch: Every mesh needs at least one triangle.");
        if (g.vertex_count == 0)
            throw_invalid("bvh.build: Cannot build a BVH with triangles but no vertices.");
        if (T > (1ull << 30) - 1ull)
            throw_invalid("bvh.build: Triangle count exceeds the native wide-collapse counter range.");
        leaf_prefix[m] = static_cast<std::uint32_t>(leaves);
        internal_prefix[m] = static_cast<std::uint32_t>(leaves - m);
        block_prefix[m] = static_cast<std::uint32_t>(blocks);
        seg_mesh sm{};
        sm.x = g.x.ptr;
        sm.y = g.y.ptr;
        sm.z = g.z.ptr;
        sm.i0 = g.i0.ptr;
        sm.i1 = g.i1.ptr;
        sm.i2 = g.i2.ptr;
        sm.vertex_count = g.vertex_count;
        sm.leaf_base = static_cast<std::uint32_t>(leaves);
        sm.leaves = static_cast<std::uint32_t>(T);
        sm.internal_base = static_cast<std::uint32_t>(leaves - m);
        t.host_meshes.push_back(sm);
        leaves += T;
        blocks += std::max<std::uint64_t>(1, std::min<std::uint64_t>(512, (g.vertex_count + k_block - 1) / k_block));
        status_words += (T - 1 + k_block - 1) / k_block;
        // Leaf positions and cluster ids (total + internal index) must fit 32 bits.
        if (leaves >= (1ull << 31) || blocks >= (1ull << 31))
            throw_invalid("batch: The meshes' triangles exceed the segmented build's 2^31 - 1 range.");
    }
    leaf_prefix[M] = static_cast<std::uint32_t>(leaves);
    internal_prefix[M] = static_cast<std::uint32_t>(leaves - M);
    block_prefix[M] = static_cast<std::uint32_t>(blocks);
    t.total = static_cast<std::uint32_t>(leaves);
    t.internal_total = static_cast<std::uint32_t>(leaves - M);
    t.vertex_blocks = static_cast<std::uint32_t>(blocks);
    t.status_words = static_cast<std::uint32_t>(std::max<std::uint64_t>(1, status_words));
    t.mesh_bits = 0;
    while ((std::uint64_t(1) << t.mesh_bits) < M)
        ++t.mesh_bits;
    int dev = 0, cus = 1;
    GWN_HIP(hipGetDevice(&dev));
    GWN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    t.grid = static_cast<unsigned>(std::max(1, cus) * 8);

    std::size_t const n = t.total, ni = t.internal_total;
    std::size_t &b = t.bytes;
    alloc(t.meshes, M, stream, b);
    alloc(t.leaf_prefix, M + 1, stream, b);
    alloc(t.block_prefix, M + 1, stream, b);
    alloc(t.internal_prefix, M + 1, stream, b);
    alloc(t.scene_words, 6 * M, stream, b);
    alloc(t.scene, 9 * M, stream, b);
    alloc(t.boxes, n, stream, b);
    alloc(t.leaf_boxes, n, stream, b);
    alloc(t.codes, n, stream, b);
    alloc(t.codes_alt, n, stream, b);
    alloc(t.ids, n, stream, b);
    alloc(t.ids_alt, n, stream, b);
    alloc(t.keys, n, stream, b);
    alloc(t.keys_alt, n, stream, b);
    alloc(t.pos, n, stream, b);
    alloc(t.pos_alt, n, stream, b);
    alloc(t.sort_ws, radix_sort_workspace_words(n, 0, 63), stream, b);
    alloc(t.bnodes, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.ibounds, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.ifirst, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.clusters, n, stream, b);
    alloc(t.boundary, n, stream, b);
    alloc(t.root_words, M, stream, b);
    alloc(t.words, 8, stream, b);
    alloc(t.parked, hploc_parked_capacity(n, k_hploc_seg_chunk_block), stream, b);
    alloc(t.iparent, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.depth, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.jump, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.by_depth, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.buckets, 3 * k_batch_build_max_levels + 1, stream, b);
    alloc(t.tasks, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.levels, M * k_level_stride, stream, b);
    alloc(t.stats, M, stream, b);
    alloc(t.status, 2 * static_cast<std::size_t>(t.status_words), stream, b);
    alloc(t.tile_prefix, M + 1, stream, b);
    alloc(t.claim, 1, stream, b);
    alloc(t.trees, M, stream, b);
    alloc(t.nodes, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.parent, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.internal_children, std::max<std::size_t>(1, ni), stream, b);
    alloc(t.prim, n, stream, b);
    alloc(t.tris, n, stream, b);
    alloc(t.root_box, M, stream, b);
    std::size_t const RV = static_cast<std::size_t>(moment_record_vec4(order));
    std::size_t const EV = static_cast<std::size_t>(eval_record_vec4(order));
    std::size_t const PB = moment_payload_bytes(order);
    alloc(t.records, std::max<std::size_t>(1, ni) * RV, stream, b);
    alloc(t.eval_records, std::max<std::size_t>(1, ni) * EV, stream, b);
    alloc(t.slot_area, std::max<std::size_t>(1, ni) * 4, stream, b);
    alloc(t.pending, std::max<std::size_t>(1, ni) * PB, stream, b);
    alloc(t.refit_meshes, M, stream, b);
    alloc(t.segments, M * static_cast<std::size_t>(k_batch_build_max_levels), stream, b);
    alloc(t.refit_levels, k_batch_build_max_levels, stream, b);
    alloc(t.slot_prefix, M + 1, stream, b);

    std::vector<collapse_tree> trees(M);
    std::vector<moment_refit_mesh> refits(M);
    for (std::size_t m = 0; m < M; ++m) {
        seg_mesh const &sm = t.host_meshes[m];
        std::size_t const lb = sm.leaf_base, ib = sm.internal_base;
        std::uint32_t const cap = sm.leaves - 1u;
        trees[m] = collapse_tree{t.bnodes.ptr + ib, t.ibounds.ptr + ib, t.ifirst.ptr + ib, t.leaf_boxes.ptr + lb,
                                 t.levels.ptr + m * k_level_stride, t.tasks.ptr + ib, t.stats.ptr + m,
                                 t.nodes.ptr + ib, t.parent.ptr + ib, t.internal_children.ptr + ib, cap};
        refits[m] = moment_refit_mesh{t.nodes.ptr + ib, t.tris.ptr + lb, t.pending.ptr + ib * PB,
                                      t.records.ptr + ib * RV, t.eval_records.ptr + ib * EV, t.slot_area.ptr + ib * 4,
                                      cap, sm.leaves};
    }
    GWN_HIP(hipMemcpyAsync(t.meshes.ptr, t.host_meshes.data(), M * sizeof(seg_mesh), hipMemcpyHostToDevice, stream));
    GWN_HIP(hipMemcpyAsync(t.leaf_prefix.ptr, leaf_prefix.data(), (M + 1) * sizeof(std::uint32_t),
                           hipMemcpyHostToDevice, stream));
    GWN_HIP(hipMemcpyAsync(t.block_prefix.ptr, block_prefix.data(), (M + 1) * sizeof(std::uint32_t),
                           hipMemcpyHostToDevice, stream));
    GWN_HIP(hipMemcpyAsync(t.internal_prefix.ptr, internal_prefix.data(), (M + 1) * sizeof(std::uint32_t),
                           hipMemcpyHostToDevice, stream));
    GWN_HIP(hipMemcpyAsync(t.trees.ptr, trees.data(), M * sizeof(collapse_tree), hipMemcpyHostToDevice, stream));
    GWN_HIP(hipMemcpyAsync(t.refit_meshes.ptr, refits.data(), M * sizeof(moment_refit_mesh), hipMemcpyHostToDevice,
                           stream));
    GWN_HIP(hipStreamSynchronize(stream));
    return tp;
}

void batch_trees_set_build(batch_trees &t, int method, int width, std::uint32_t hploc_radius) {
    if (method != GWN_BVH_BUILD_HPLOC && method != GWN_BVH_BUILD_LBVH)
        throw_invalid("bvh.build: BVH build method is invalid.");
    if (width < 2 || width > 4)
        throw_invalid("bvh.build: BVH width must be 2, 3 or 4 (nodes are stored four-wide).");
    if (method == GWN_BVH_BUILD_HPLOC && (hploc_radius == 0 || hploc_radius > 8))
        throw_invalid("bvh.build: H-PLOC search radius must be in the range [1, 8].");
    t.method = method;
    t.width = width;
    t.radius = hploc_radius;
}

void batch_trees_enqueue(batch_trees &t, int levels, int binary_levels, std::uint32_t stack_capacity,
                         std::uint32_t *status, hipStream_t stream) {
    std::uint32_t const M = static_cast<std::uint32_t>(t.mesh_count);
    if (M == 0)
        return;
    if (levels < 1 || levels > k_batch_build_max_levels || binary_levels < 1 || binary_levels > k_batch_build_max_levels)
        throw_invalid("batch: Level launch counts must be in [1, 128].");
    std::uint32_t const n = t.total;
    // Scene bounds per mesh.
    seg_scene_init_kernel<<<blocks_for(M, k_block), k_block, 0, stream>>>(t.scene_words.ptr, M);
    GWN_LAUNCH_CHECK();
    seg_scene_bounds_kernel<<<t.vertex_blocks, k_block, 0, stream>>>(t.meshes.ptr, t.block_prefix.ptr, M,
                                                                      t.scene_words.ptr);
    GWN_LAUNCH_CHECK();
    seg_scene_finalize_kernel<<<blocks_for(M, k_block), k_block, 0, stream>>>(t.scene_words.ptr, t.scene.ptr, M);
    GWN_LAUNCH_CHECK();
    // Codes, then the stable sorts by code and by mesh.
    seg_morton_kernel<<<blocks_for(n, k_block), k_block, 0, stream>>>(t.meshes.ptr, t.leaf_prefix.ptr, M, t.scene.ptr,
                                                                      n, t.boxes.ptr, t.codes.ptr, t.ids.ptr);
    GWN_LAUNCH_CHECK();
    std::uint64_t *codes1 = nullptr;
    std::uint32_t *ids1 = nullptr;
    radix_sort_pairs_u64_ws(t.codes.ptr, t.codes_alt.ptr, t.ids.ptr, t.ids_alt.ptr, n, 0, 63, t.sort_ws.ptr, &codes1,
                            &ids1, stream);
    std::uint32_t *mesh_of = nullptr, *pos_of = nullptr;
    if (M > 1) {
        seg_mesh_key_kernel<<<blocks_for(n, k_block), k_block, 0, stream>>>(ids1, t.leaf_prefix.ptr, M, n, t.keys.ptr,
                                                                            t.pos.ptr);
        GWN_LAUNCH_CHECK();
        radix_sort_pairs_u32_ws(t.keys.ptr, t.keys_alt.ptr, t.pos.ptr, t.pos_alt.ptr, n, 0, t.mesh_bits, t.sort_ws.ptr,
                                &mesh_of, &pos_of, stream);
    }
    std::uint64_t *codes = codes1 == t.codes.ptr ? t.codes_alt.ptr : t.codes.ptr;
    fill_u32(t.words.ptr, 0u, 8, stream);
    seg_gather_kernel<<<blocks_for(n, k_block), k_block, 0, stream>>>(
        t.meshes.ptr, n, mesh_of, pos_of, codes1, ids1, t.boxes.ptr, codes, t.leaf_boxes.ptr, t.prim.ptr, t.tris.ptr,
        t.words.ptr + 2);
    GWN_LAUNCH_CHECK();
    std::size_t const ni = t.internal_total;
    std::uint32_t const NI = t.internal_total;
    if (t.method == GWN_BVH_BUILD_LBVH) {
        // Karras LBVH, binary depths by pointer jumping, node ids bucketed by depth, then bounds one
        // binary depth per launch, deepest first.
        fill_u32(t.root_words.ptr, 0u, M, stream);
        if (ni > 0) {
            seg_lbvh_kernel<<<blocks_for(ni, k_block), k_block, 0, stream>>>(t.meshes.ptr, t.internal_prefix.ptr, M, NI,
                                                                             codes, t.bnodes.ptr, t.iparent.ptr);
            GWN_LAUNCH_CHECK();
            unsigned const nb = blocks_for(ni, k_block);
            seg_jump_init_kernel<<<nb, k_block, 0, stream>>>(t.meshes.ptr, t.internal_prefix.ptr, M, NI, t.iparent.ptr,
                                                            t.jump.ptr, t.words.ptr + 1);
            GWN_LAUNCH_CHECK();
            for (int r = 0; r < k_jump_rounds; ++r) {
                seg_jump_kernel<<<nb, k_block, 0, stream>>>(NI, t.jump.ptr);
                GWN_LAUNCH_CHECK();
            }
            std::uint32_t *counts = t.buckets.ptr;
            std::uint32_t *starts = t.buckets.ptr + k_batch_build_max_levels;
            std::uint32_t *cursor = t.buckets.ptr + 2 * k_batch_build_max_levels + 1;
            fill_u32(counts, 0u, k_batch_build_max_levels, stream);
            seg_depth_kernel<<<t.grid, k_block, 0, stream>>>(NI, t.jump.ptr, t.depth.ptr, counts, t.words.ptr + 3);
            GWN_LAUNCH_CHECK();
            seg_depth_scan_kernel<<<1, k_batch_build_max_levels, 0, stream>>>(counts, starts, cursor);
            GWN_LAUNCH_CHECK();
            seg_depth_scatter_kernel<<<t.grid, k_block, 0, stream>>>(NI, t.depth.ptr, starts, cursor, t.by_depth.ptr);
            GWN_LAUNCH_CHECK();
            for (int d = binary_levels; d-- > 0;) {
                seg_lbvh_bounds_level_kernel<<<t.grid, k_block, 0, stream>>>(
                    t.meshes.ptr, t.internal_prefix.ptr, M, t.by_depth.ptr, starts, static_cast<std::uint32_t>(d),
                    t.leaf_boxes.ptr, t.bnodes.ptr, t.ibounds.ptr, t.ifirst.ptr);
                GWN_LAUNCH_CHECK();
            }
        }
    } else {
        // H-PLOC over every mesh.
        if (ni > 0) {
            fill_u32(reinterpret_cast<std::uint32_t *>(t.bnodes.ptr), k_leaf_bit, 2 * ni, stream);
            fill_u32(t.ifirst.ptr, 0u, ni, stream);
        }
        fill_u32(t.root_words.ptr, k_invalid, M, stream);
        seg_hploc_init_kernel<<<blocks_for(n, k_block), k_block, 0, stream>>>(t.clusters.ptr, t.boundary.ptr, n);
        GWN_LAUNCH_CHECK();
        hploc_out const out_arrays{t.bnodes.ptr, t.ibounds.ptr, t.ifirst.ptr, t.clusters.ptr};
        seg_hploc_kernel<<<blocks_for(n, k_hploc_seg_chunk_block), k_hploc_seg_chunk_block, 0, stream>>>(
            t.meshes.ptr, t.leaf_prefix.ptr, M, codes, t.leaf_boxes.ptr, n, t.radius, out_arrays, t.boundary.ptr,
            t.parked.ptr, t.root_words.ptr, t.words.ptr);
        GWN_LAUNCH_CHECK();
        seg_hploc_parked_kernel<<<blocks_for(t.parked.count, k_hploc_block), k_hploc_block, 0, stream>>>(
            t.meshes.ptr, t.leaf_prefix.ptr, M, codes, t.leaf_boxes.ptr, n, t.radius, out_arrays, t.boundary.ptr,
            t.parked.ptr, t.root_words.ptr, t.words.ptr);
        GWN_LAUNCH_CHECK();
    }
    // Collapse, one plan + one persistent launch per wide level.
    fill_u32(reinterpret_cast<std::uint32_t *>(t.levels.ptr), 0u, static_cast<std::size_t>(M) * k_level_stride * 4,
             stream);
    fill_u32(t.status.ptr, 0u, 2 * static_cast<std::size_t>(t.status_words), stream);
    seg_collapse_init_kernel<<<blocks_for(M, k_block), k_block, 0, stream>>>(t.meshes.ptr, t.trees.ptr, M,
                                                                             t.root_words.ptr, t.root_box.ptr);
    GWN_LAUNCH_CHECK();
    unsigned const collapse_grid = std::min<unsigned>(t.grid, std::max<unsigned>(1, t.status_words));
    for (int d = 0; d < levels; ++d) {
        std::uint32_t *cur = t.status.ptr + static_cast<std::size_t>(d & 1) * t.status_words;
        std::uint32_t *nxt = t.status.ptr + static_cast<std::size_t>((d + 1) & 1) * t.status_words;
        seg_collapse_plan_kernel<<<1, k_plan_block, 0, stream>>>(t.trees.ptr, M, d, t.tile_prefix.ptr, t.claim.ptr);
        GWN_LAUNCH_CHECK();
        seg_collapse_level_kernel<<<collapse_grid, k_block, 0, stream>>>(t.trees.ptr, M, d, t.tile_prefix.ptr,
                                                                         t.claim.ptr, cur, nxt, t.status_words,
                                                                         t.width);
        GWN_LAUNCH_CHECK();
    }
    std::uint32_t const merges = t.method == GWN_BVH_BUILD_HPLOC ? NI : 0u;
    std::uint32_t const bin_check = t.method == GWN_BVH_BUILD_LBVH ? static_cast<std::uint32_t>(binary_levels) : ~0u;
    seg_check_kernel<<<blocks_for(M, k_block), k_block, 0, stream>>>(t.meshes.ptr, t.trees.ptr, M, levels,
                                                                     stack_capacity, t.words.ptr, merges, bin_check,
                                                                     status);
    GWN_LAUNCH_CHECK();
    // Moments.
    seg_refit_plan_kernel<<<levels + 1, k_plan_block, 0, stream>>>(t.meshes.ptr, t.trees.ptr, M, levels,
                                                                   t.segments.ptr, t.refit_levels.ptr,
                                                                   t.slot_prefix.ptr);
    GWN_LAUNCH_CHECK();
    refit_moment_batch_device(t.refit_meshes.ptr, t.segments.ptr, t.refit_levels.ptr, levels, t.slot_prefix.ptr, M,
                              t.order, status, t.grid, stream);
}

batch_trees_summary batch_trees_read(batch_trees &t, hipStream_t stream) {
    batch_trees_summary s;
    std::size_t const M = t.mesh_count;
    if (M == 0)
        return s;
    std::vector<collapse_level> lv(M * k_level_stride);
    std::vector<collapse_stats> st(M);
    s.root_boxes.resize(M);
    GWN_HIP(hipMemcpyAsync(lv.data(), t.levels.ptr, lv.size() * sizeof(collapse_level), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(st.data(), t.stats.ptr, M * sizeof(collapse_stats), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(s.root_boxes.data(), t.root_box.ptr, M * sizeof(aabb), hipMemcpyDeviceToHost, stream));
    std::uint32_t words[4] = {0, 0, 0, 0};
    GWN_HIP(hipMemcpyAsync(words, t.words.ptr, sizeof(words), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    s.max_binary_depth = t.method == GWN_BVH_BUILD_LBVH ? words[3] : 0u;
    for (std::size_t m = 0; m < M; ++m) {
        if (t.host_meshes[m].leaves < 2u)
            continue;
        for (std::uint32_t d = 0; d < k_level_stride; ++d)
            if (lv[m * k_level_stride + d].count != 0u)
                s.max_depth = std::max(s.max_depth, d);
        s.max_internal_stack = std::max(s.max_internal_stack, st[m].max_internal);
    }
    return s;
}

void batch_trees_view(batch_trees const &t, std::size_t i, float4 const **eval_records, tri_record const **tris,
                      std::uint64_t *root) {
    seg_mesh const &sm = t.host_meshes[i];
    bool const internal = sm.leaves >= 2u;
    *eval_records = internal ? t.eval_records.ptr + static_cast<std::size_t>(sm.internal_base) *
                                                        static_cast<std::size_t>(eval_record_vec4(t.order))
                             : nullptr;
    *tris = t.tris.ptr + sm.leaf_base;
    *root = internal ? encode_internal(0) : encode_leaf(0, 1);
}

void batch_trees_export(batch_trees &t, std::size_t i, gwn_bvh_handle &bvh, gwn_moment_handle &moment,
                        hipStream_t stream) {
    if (i >= t.mesh_count)
        throw_invalid("batch: Mesh index is out of range.");
    seg_mesh const &sm = t.host_meshes[i];
    std::vector<collapse_level> lv(k_level_stride);
    collapse_stats st{};
    aabb box{};
    GWN_HIP(hipMemcpyAsync(lv.data(), t.levels.ptr + i * k_level_stride, k_level_stride * sizeof(collapse_level),
                           hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&st, t.stats.ptr + i, sizeof(st), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&box, t.root_box.ptr + i, sizeof(box), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    std::size_t const T = sm.leaves, lb = sm.leaf_base, ib = sm.internal_base;
    gwn_bvh_handle b;
    b.stream = stream;
    b.triangle_count = T;
    b.width = t.width;
    b.primitive_indices.allocate(T, stream);
    b.triangles.allocate(T, stream);
    GWN_HIP(hipMemcpyAsync(b.primitive_indices.ptr, t.prim.ptr + lb, T * sizeof(std::uint32_t),
                           hipMemcpyDeviceToDevice, stream));
    GWN_HIP(hipMemcpyAsync(b.triangles.ptr, t.tris.ptr + lb, T * sizeof(tri_record), hipMemcpyDeviceToDevice, stream));
    b.root.bounds = box;
    gwn_moment_handle mo;
    mo.order = t.order;
    mo.width = t.width;
    mo.stream = stream;
    if (T < 2) {
        b.root.reference = encode_leaf(0, 1);
    } else {
        if (st.error != 0u || st.leaves != T)
            throw_internal("batch: The segmented build of this mesh failed.");
        int empty = -1;
        for (int d = 0; d < static_cast<int>(k_level_stride); ++d)
            if (lv[d].count == 0u) {
                empty = d;
                break;
            }
        if (empty <= 0)
            throw_internal("batch: The segmented build's level table of this mesh is incomplete.");
        std::size_t const NW = lv[empty].start;
        b.nodes.allocate(NW, stream);
        b.parent.allocate(NW, stream);
        b.internal_children.allocate(NW, stream);
        GWN_HIP(hipMemcpyAsync(b.nodes.ptr, t.nodes.ptr + ib, NW * sizeof(bvh_node4), hipMemcpyDeviceToDevice, stream));
        GWN_HIP(hipMemcpyAsync(b.parent.ptr, t.parent.ptr + ib, NW * sizeof(std::uint32_t), hipMemcpyDeviceToDevice,
                               stream));
        GWN_HIP(hipMemcpyAsync(b.internal_children.ptr, t.internal_children.ptr + ib, NW, hipMemcpyDeviceToDevice,
                               stream));
        b.node_count = NW;
        b.root.reference = encode_internal(0);
        b.internal_stack_bound = st.max_internal;
        b.packed_stack_bound = st.max_packed;
        b.max_depth = static_cast<std::uint32_t>(empty - 1);
        b.level_start.resize(static_cast<std::size_t>(empty) + 1);
        for (int d = 0; d <= empty; ++d)
            b.level_start[d] = lv[d].start;
        std::size_t const RV = static_cast<std::size_t>(moment_record_vec4(t.order));
        std::size_t const EV = static_cast<std::size_t>(eval_record_vec4(t.order));
        mo.records.allocate(NW * RV, stream);
        mo.eval_records.allocate(NW * EV, stream);
        mo.slot_area.allocate(NW * 4, stream);
        GWN_HIP(hipMemcpyAsync(mo.records.ptr, t.records.ptr + ib * RV, NW * RV * sizeof(float4),
                               hipMemcpyDeviceToDevice, stream));
        GWN_HIP(hipMemcpyAsync(mo.eval_records.ptr, t.eval_records.ptr + ib * EV, NW * EV * sizeof(float4),
                               hipMemcpyDeviceToDevice, stream));
        GWN_HIP(hipMemcpyAsync(mo.slot_area.ptr, t.slot_area.ptr + ib * 4, NW * 4 * sizeof(float),
                               hipMemcpyDeviceToDevice, stream));
        mo.node_count = NW;
    }
    GWN_HIP(hipStreamSynchronize(stream));
    b.storage_generation = next_storage_generation();
    b.revision = reserve_revision();
    if (b.revision == 0)
        throw_internal("bvh.build: BVH geometry-derived state revision is exhausted.");
    mo.bvh_revision = b.revision;
    mo.storage_generation = next_storage_generation();
    std::swap(bvh, b);
    std::swap(moment, mo);
}

int batch_trees_width(batch_trees const &t) { return t.width; }
int batch_trees_method(batch_trees const &t) { return t.method; }
std::uint32_t batch_trees_radius(batch_trees const &t) { return t.radius; }
std::size_t batch_trees_bytes(batch_trees const &t) { return t.bytes; }

} // namespace gwn
