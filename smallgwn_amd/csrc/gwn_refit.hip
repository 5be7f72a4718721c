// gwn_refit.hip - in-place bounds refit of a canonical BVH4 from updated geometry (gfx950).
//
// Reference: include/gwn/detail/gwn_bvh_refit_impl.cuh:23-352 (from-leaves climb, finalize,
// leaf-root worker).  Topology, primitive order and both stack bounds are kept; the
// leaf-ordered triangle records and every child box are rewritten.  A node's box is the fold of
// its child boxes in slot order, bounds = union(bounds, child) (combine_node, :40-51), and leaf
// boxes come from the same vertex loads as the records (gwn_bvh_triangle_impl.cuh:56-66).
//
// One launch per breadth-first level of the wide tree, deepest first.  A thread takes one node of
// the level: it rewrites the records of its leaf slots, writes every child box (a leaf's from its
// triangles, an internal child's from the fold the previous, deeper launch stored) and stores its
// own fold for its parent's level.  The kernel boundary between levels orders the passes: no
// arrival counters or cross-workgroup fences.  Child boxes are written straight into the node
// array (the reference stages them and commits after convergence; the BVH is marked unqueryable
// for the duration in both designs).
#include "gwn_core.hpp"

// Box folds must match the oracle bit for bit; there is no arithmetic to contract here, but keep
// the build's convention.
#pragma clang fp contract(off)

namespace gwn {
namespace {

constexpr int k_block = 256;

__device__ inline float min_ref(float a, float b) { return (b < a) ? b : a; } // std::min(a, b)
__device__ inline float max_ref(float a, float b) { return (a < b) ? b : a; } // std::max(a, b)

__device__ inline aabb union_ref(aabb const &a, aabb const &b) {
    return aabb{min_ref(a.min_x, b.min_x), min_ref(a.min_y, b.min_y), min_ref(a.min_z, b.min_z),
                max_ref(a.max_x, b.max_x), max_ref(a.max_y, b.max_y), max_ref(a.max_z, b.max_z)};
}

struct refit_geometry {
    float const *x, *y, *z;
    std::uint32_t const *i0, *i1, *i2;
    std::size_t vertex_count, triangle_count;
    std::uint32_t const *primitive_indices;
    tri_record *triangles;
};

// Rewrites the records of one leaf range and returns its box; false on an invalid index.
__device__ inline bool refit_leaf_range(
    refit_geometry const &g, std::uint64_t offset, std::uint32_t count, aabb &box
) {
    if (count == 0u || offset + count > g.triangle_count)
        return false;
    for (std::uint32_t j = 0; j < count; ++j) {
        std::uint32_t const p = g.primitive_indices[offset + j];
        if (p >= g.triangle_count)
            return false;
        std::uint32_t const a = g.i0[p], b = g.i1[p], c = g.i2[p];
        if (a >= g.vertex_count || b >= g.vertex_count || c >= g.vertex_count)
            return false;
        float const ax = g.x[a], ay = g.y[a], az = g.z[a];
        float const bx = g.x[b], by = g.y[b], bz = g.z[b];
        float const cx = g.x[c], cy = g.y[c], cz = g.z[c];
        g.triangles[offset + j] = tri_record{ax, ay, az, bx, by, bz, cx, cy, cz};
        aabb const t{min_ref(ax, min_ref(bx, cx)), min_ref(ay, min_ref(by, cy)), min_ref(az, min_ref(bz, cz)),
                     max_ref(ax, max_ref(bx, cx)), max_ref(ay, max_ref(by, cy)), max_ref(az, max_ref(bz, cz))};
        box = j == 0 ? t : union_ref(box, t);
    }
    return true;
}

__device__ inline void refit_node(refit_geometry const &g, bvh_node4 *nodes, std::size_t node_count, std::size_t n,
                                  aabb *pending, std::uint32_t *error) {
    aabb acc{};
    bool init = false;
    for (int s = 0; s < 4; ++s) {
        std::uint64_t const ref = nodes[n].children[s].reference;
        if (!ref_valid(ref))
            continue;
        aabb box;
        if (ref_internal(ref)) {
            std::uint64_t const c = ref_offset(ref);
            if (c >= node_count || c <= n) {
                atomicOr(error, 1u);
                return;
            }
            box = pending[c];
        } else if (!refit_leaf_range(g, ref_offset(ref), ref_count(ref), box)) {
            atomicOr(error, 1u);
            return;
        }
        nodes[n].children[s].bounds = box;
        acc = init ? union_ref(acc, box) : box;
        init = true;
    }
    if (!init) {
        atomicOr(error, 1u);
        return;
    }
    pending[n] = acc;
}

__global__ __launch_bounds__(k_block) void refit_level_kernel(
    refit_geometry g, bvh_node4 *nodes, std::size_t node_count, std::size_t first, std::size_t count,
    aabb *pending, std::uint32_t *error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i < count)
        refit_node(g, nodes, node_count, first + i, pending, error);
}

// A run of small levels (at most k_fused_block nodes each) in one workgroup, a barrier between
// levels.
constexpr int k_fused_block = 1024;
__global__ __launch_bounds__(k_fused_block) void refit_levels_fused_kernel(
    refit_geometry g, bvh_node4 *nodes, std::size_t node_count, level_run run, aabb *pending, std::uint32_t *error
) {
    for (int k = 0; k < run.levels; ++k) {
        if (threadIdx.x < run.count[k])
            refit_node(g, nodes, node_count, run.first[k] + threadIdx.x, pending, error);
        __syncthreads();
    }
}

// A leaf root owns the whole (at most 65535-primitive) range: one thread, no climb.
__global__ void refit_leaf_root_kernel(refit_geometry g, std::uint64_t root_ref, aabb *root_box, std::uint32_t *error) {
    aabb box;
    if (!refit_leaf_range(g, ref_offset(root_ref), ref_count(root_ref), box)) {
        atomicOr(error, 1u);
        return;
    }
    *root_box = box;
}

} // namespace

void refit_bounds(gwn_geometry_handle const &geo, gwn_bvh_handle &bvh, hipStream_t stream) {
    std::size_t const N = bvh.node_count;
    refit_geometry const g{geo.x.ptr, geo.y.ptr, geo.z.ptr, geo.i0.ptr, geo.i1.ptr, geo.i2.ptr,
                           geo.vertex_count, geo.triangle_count, bvh.primitive_indices.ptr,
                           bvh.triangles.ptr};
    device_buffer<std::uint32_t> error(1, stream);
    fill_u32(error.ptr, 0u, 1, stream);
    device_buffer<aabb> boxes(N > 0 ? N : 1, stream);
    if (ref_internal(bvh.root.reference)) {
        auto const &ls = bvh.level_start;
        if (ls.size() < 2 || ls.front() != 0u || ls.back() != N)
            throw_internal("bvh.refit: BVH level table does not cover its nodes.");
        for (level_run const &r : plan_level_runs(ls, k_fused_block)) {
            if (r.levels > 0)
                refit_levels_fused_kernel<<<1, k_fused_block, 0, stream>>>(g, bvh.nodes.ptr, N, r, boxes.ptr,
                                                                           error.ptr);
            else
                refit_level_kernel<<<grid_for(r.big_count, k_block), k_block, 0, stream>>>(
                    g, bvh.nodes.ptr, N, r.big_first, r.big_count, boxes.ptr, error.ptr);
            GWN_LAUNCH_CHECK();
        }
    } else {
        refit_leaf_root_kernel<<<1, 1, 0, stream>>>(g, bvh.root.reference, boxes.ptr, error.ptr);
        GWN_LAUNCH_CHECK();
    }
    std::uint32_t host_error = 0;
    aabb root_box{};
    GWN_HIP(hipMemcpyAsync(&host_error, error.ptr, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&root_box, boxes.ptr, sizeof(aabb), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.refit: Bottom-up BVH refit did not converge.");
    bvh.root.bounds = root_box;
}

} // namespace gwn
