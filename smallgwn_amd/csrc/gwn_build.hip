// gwn_build.hip - canonical BVH4 construction on gfx950.
//
// Pipeline (reference: include/gwn/detail/gwn_bvh_build_impl.cuh:44-255):
//   scene bounds (device reduction, no host read-back)       build_common.cuh:243-274
//   triangle AABB + 63-bit Morton code                       build_common.cuh:157-214
//   stable LSD radix sort of (code, triangle id)             build_common.cuh:325-347
//   Karras LBVH                                              build_binary.cuh:29-129
//   bounds + first leaf, one launch per binary depth         build_binary.cuh:131-192
//     (or H-PLOC, which writes them beside each node)        build_hploc.cuh:104-622
//   BVH4 collapse, one launch per level, emitting nodes      build_collapse.cuh:98-314
//     directly in breadth-first (depth, first sorted leaf)   (replaces reorder.cuh:61-143)
//     order
//   leaf-ordered triangle records                            build_impl.cuh:23-41
// The host reads the small level table back once per chunk of levels (about three times for a
// million triangles) to find the last level and size the node array; the reference performs
// one such read-back of its node count (collapse.cuh:425-446).  The device pieces shared with the
// segmented build of the batched mode are in gwn_build_device.hpp.
#include "gwn_build_device.hpp"

#include <algorithm>
#include <vector>

namespace gwn {
namespace {

// Per-axis min/max over all vertices; one atomic per block per extremum.
__global__ __launch_bounds__(k_block) void scene_bounds_kernel(
    float const *__restrict__ x, float const *__restrict__ y, float const *__restrict__ z,
    std::size_t nv, std::uint32_t *__restrict__ ordered
) {
    scene_bounds_range(x, y, z, static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x, nv,
                       static_cast<std::size_t>(gridDim.x) * k_block, ordered);
}

__global__ void scene_init_kernel(std::uint32_t *ordered) { scene_init_words(ordered); }

__global__ void scene_finalize_kernel(std::uint32_t const *ordered, float *scene) {
    scene_finalize_words(ordered, scene);
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
    aabb box;
    std::uint64_t code;
    morton_triangle(vx, vy, vz, nv, i0, i1, i2, t, scene, box, code);
    boxes[t] = box;
    codes[t] = code;
}


__global__ __launch_bounds__(k_block) void gather_boxes_kernel(
    aabb const *__restrict__ boxes, std::uint32_t const *__restrict__ order, std::size_t nt,
    aabb *__restrict__ sorted
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < nt)
        sorted[i] = boxes[order[i]];
}

// Karras LBVH over one tree, one thread per internal node (lbvh_node).
__global__ __launch_bounds__(k_block) void lbvh_kernel(
    std::uint64_t const *__restrict__ codes, std::size_t n, uint2 *__restrict__ bnodes,
    std::uint32_t *__restrict__ iparent, std::uint32_t *__restrict__ lparent
) {
    std::size_t const tid = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (tid + 1 >= n)
        return;
    lbvh_node(codes, static_cast<std::int64_t>(n), static_cast<std::int64_t>(tid), bnodes, iparent, lparent);
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
        d = lbvh_node_depth(iparent, ni, i, error);
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
    lbvh_node_bounds(leaf_boxes, bnodes, node, ibounds, ifirst);
}

// H-PLOC over one tree: the chunk launch, one lane per sorted leaf, then the global launch over
// the ranges it parked (hploc_chunk_climb, hploc_climb).  words: [0] merges, [1] failure flag,
// [2] root internal id, [3] parked ranges.
__global__ __launch_bounds__(k_hploc_chunk_block) void hploc_kernel(
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t T,
    std::uint32_t radius, hploc_out o, std::uint32_t *boundary, uint2 *parked, std::uint32_t *words
) {
    std::uint32_t const idx = blockIdx.x * k_hploc_chunk_block + threadIdx.x;
    hploc_chunk_climb<k_hploc_chunk_block>(codes, leaf_boxes, T, radius, o, boundary, build_segment{0u, T, 0u, 0u},
                                           idx, idx < T, parked, words + 3, words + 2, words);
}

__global__ __launch_bounds__(k_hploc_block) void hploc_parked_kernel(
    std::uint64_t const *__restrict__ codes, aabb const *__restrict__ leaf_boxes, std::uint32_t T,
    std::uint32_t radius, hploc_out o, std::uint32_t *boundary, uint2 const *parked, std::uint32_t *words
) {
    std::uint32_t const i = hploc_parked_slot(blockIdx.x);
    bool const active = i != k_invalid && i < words[3];
    uint2 const r = active ? parked[i] : make_uint2(0u, 0u);
    hploc_climb<false>(codes, leaf_boxes, T, radius, o, boundary, build_segment{0u, T, 0u, 0u}, r.x, r.y, active, 0u,
                       0u, nullptr, nullptr, words + 2, words);
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

// One launch per level of the collapse (collapse_tile).
__global__ __launch_bounds__(k_block) void collapse_level_kernel(
    collapse_tree t, int depth, std::uint32_t *status_cur, std::uint32_t *status_next, std::uint32_t status_words,
    int width
) {
    __shared__ std::uint32_t s_tile;
    // Only the first ceil(count / block) blocks take part (the launch is sized for the level's
    // bound, not its count): they claim exactly the tiles 0..tiles-1, each claim made by a running
    // block, so the look-back never waits on an unclaimed tile.  A level past the last has count 0
    // (the level table starts zeroed) and every block returns at once; no launch reads the next
    // status array then.
    std::uint64_t const level_tiles = (static_cast<std::uint64_t>(t.levels[depth].count) + k_block - 1) / k_block;
    if (blockIdx.x >= level_tiles)
        return;
    std::uint32_t const live = level_tiles < gridDim.x ? static_cast<std::uint32_t>(level_tiles) : gridDim.x;
    if (threadIdx.x == 0)
        s_tile = atomicAdd(&t.levels[depth].tiles_claimed, 1u);
    // The next level scans over the other status array; the live blocks clear it for that launch.
    for (std::uint32_t i = blockIdx.x * k_block + threadIdx.x; i < status_words; i += live * k_block)
        status_next[i] = 0u;
    __syncthreads();
    std::uint32_t const start = t.levels[depth].start;
    std::uint32_t const count = t.levels[depth].count;
    std::uint32_t const tile = s_tile;
    // A level never exceeds the launch (width^depth bound) or the capacity; report, do not index.
    if (static_cast<std::uint64_t>(start) + count > t.capacity ||
        count > static_cast<std::uint64_t>(gridDim.x) * k_block) {
        if (threadIdx.x == 0)
            atomicOr(&t.stats->error, 1u);
        return;
    }
    if (tile >= (count + k_block - 1) / k_block)
        return;
    collapse_tile(t, depth, start, count, tile, status_cur, width);
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
    if (!triangle_record(vx, vy, vz, nv, i0, i1, i2, order[i], records + i))
        atomicOr(error, 1u);
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
        // At most two parked ranges per chunk (one holding its first leaf, one its last).
        std::size_t const parked_cap = hploc_parked_capacity(T, k_hploc_chunk_block);
        device_buffer<uint2> parked(parked_cap, stream);
        hploc_words.allocate(4, stream);
        fill_u32(hploc_words.ptr, 0u, 4, stream);
        // Every node id is written exactly once by a correct run; the fill keeps any node a
        // failed run leaves unwritten in bounds (two references to leaf 0) for the collapse.
        fill_u32(reinterpret_cast<std::uint32_t *>(bnodes.ptr), k_leaf_bit, 2 * NI, stream);
        fill_u32(ifirst.ptr, 0u, NI, stream);
        hploc_init_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
            clusters.ptr, boundary.ptr, static_cast<std::uint32_t>(T));
        GWN_LAUNCH_CHECK();
        hploc_out const out_arrays{bnodes.ptr, ibounds.ptr, ifirst.ptr, clusters.ptr};
        hploc_kernel<<<grid_for(T, k_hploc_chunk_block), k_hploc_chunk_block, 0, stream>>>(
            sorted_codes, leaf_boxes.ptr, static_cast<std::uint32_t>(T), hploc_radius, out_arrays, boundary.ptr,
            parked.ptr, hploc_words.ptr
        );
        GWN_LAUNCH_CHECK();
        hploc_parked_kernel<<<static_cast<unsigned>(hploc_parked_blocks(parked_cap)), k_hploc_block, 0, stream>>>(
            sorted_codes, leaf_boxes.ptr, static_cast<std::uint32_t>(T), hploc_radius, out_arrays, boundary.ptr,
            parked.ptr, hploc_words.ptr
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

    collapse_tree const tree{bnodes.ptr, ibounds.ptr, ifirst.ptr, leaf_boxes.ptr, levels.ptr, tasks.ptr, stats.ptr,
                             wide.ptr, wparent.ptr, wchildren.ptr, static_cast<std::uint32_t>(capacity)};
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
            collapse_level_kernel<<<grid, k_block, 0, stream>>>(tree, depth, cur, nxt, status_words, width);
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
