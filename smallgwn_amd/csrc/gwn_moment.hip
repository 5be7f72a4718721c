// gwn_moment.hip - bottom-up Taylor moment refit (orders 0, 1, 2) on gfx950.
//
// Reference: include/gwn/detail/gwn_bvh_refit_moment_impl.cuh:17-279 and
// include/gwn/detail/gwn_bvh_refit_common.cuh:14-609.
//
// One launch per breadth-first level of the wide tree, deepest first.  A thread takes one node
// of the level; it computes the additive payload of each child slot (leaf payloads straight from
// the triangle records; an internal child's total was stored by the previous, deeper launch),
// writes the node's recentred coefficients and stores the node's total.  Children are always
// summed in slot order from the first valid slot, exactly as the reference's combine_node, and
// the kernel boundary between levels is the only ordering needed: no arrival counters, atomics or
// cross-workgroup fences.
//
// Output: one traversal record per node -- a uint4 of compact child references followed by the
// reference's field-SoA gwn_bvh_moment_node<4, Order, float> coefficients as float4 (one float
// per child slot), padded to 128 / 256 / 384 bytes -- then, in one pass over all nodes, one
// evaluation record of the same layout with the coefficients the value-only queries read
// (taylor_eval_fields).
#include "gwn_core.hpp"
#include "gwn_moment_math.hpp"

// Bit-for-bit agreement with the CPU oracle requires unfused float arithmetic.
#pragma clang fp contract(off)

namespace gwn {
namespace {

using namespace moment_math;
constexpr int k_block = 128;
constexpr std::uint32_t k_invalid = 0xffffffffu;

// One child slot of a level pass, four lanes per node (lane & 3 = the slot; the lanes of a node
// are one quad of a wave, and every lane of the wave takes part, `active` or not): the slot's
// payload (a leaf's straight from its triangle record, an internal child's from the total the
// previous, deeper level wrote), its recentred coefficients and area, then the node's total in
// slot order from the first valid slot (the reference's combine_node,
// gwn_bvh_refit_moment_impl.cuh:17-99), folded by the quad's slot-0 lane from the payloads the
// quad leaves in `share` / `share_have` (one entry per lane of the workgroup; every lane of the
// workgroup calls this, since it holds a workgroup barrier).  The whole record is written, empty slots and padding as
// zeros, so the output needs no clearing pass; the four lanes write each field's four slots as
// one 16-byte span.
template <int Order>
__device__ inline void moment_refit_slot(
    std::size_t n, int s, bool active, bvh_node4 const *__restrict__ nodes, std::size_t node_count,
    tri_record const *__restrict__ tris, std::size_t tri_count, payload<Order> *__restrict__ pending,
    float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error, payload<Order> *share,
    std::uint32_t *share_have
) {
    constexpr int RV = moment_record_vec4(Order);
    constexpr int F = moment_field_count(Order);
    payload<Order> child;
    payload_zero(child);
    bool have = false;
    if (active) {
        float4 *rec = records + n * RV;
        float *fields = reinterpret_cast<float *>(rec + 1);
        bvh_child const ch = nodes[n].children[s];
        std::uint64_t const ref = ch.reference;
        std::uint32_t cref = k_invalid;
        bool ok = true;
        if (ref_valid(ref)) {
            if (ref_internal(ref)) {
                std::uint64_t const c = ref_offset(ref);
                if (c >= node_count || c <= n) {
                    ok = false;
                } else {
                    child = pending[c];
                    cref = static_cast<std::uint32_t>(c);
                }
            } else {
                // A leaf payload starts from zero and adds each primitive in order
                // (gwn_bvh_refit_common.cuh:452-468).
                std::uint64_t const off = ref_offset(ref);
                std::uint32_t const cnt = ref_count(ref);
                if (off + cnt > tri_count || cnt != 1u) {
                    ok = false;
                } else {
                    payload<Order> tri;
                    triangle_payload<Order>(tris[off], tri);
                    payload_add(child, tri);
                    cref = k_cref_leaf | static_cast<std::uint32_t>(off);
                }
            }
            if (!ok) {
                atomicOr(error, 1u);
                payload_zero(child);
            }
        }
        have = ok && ref_valid(ref);
        if (have) {
            write_fields<Order>(fields, s, ch.bounds, child);
        } else {
            for (int f = 0; f < F; ++f)
                fields[f * 4 + s] = 0.0f;
        }
        slot_area[n * 4 + s] = have ? child.v[0] : 0.0f;
        reinterpret_cast<std::uint32_t *>(rec)[s] = cref;
        for (int k = 4 + F * 4 + s; k < RV * 4; k += 4)
            reinterpret_cast<float *>(rec)[k] = 0.0f;
    }
    // Slot order from the first valid slot: each lane leaves its payload (zero if its slot is
    // empty) and a flag, then the quad's lane 0 folds lanes 1..3 in.
    std::uint32_t const lane = threadIdx.x;
    share[lane] = child;
    share_have[lane] = have ? 1u : 0u;
    __syncthreads();
    if (active && s == 0) {
        payload<Order> total;
        bool init = have;
        if (have)
            total = share[lane];
        for (int k = 1; k < 4; ++k) {
            if (share_have[lane + k] == 0u)
                continue;
            payload<Order> const &o = share[lane + k];
            if (init) {
                payload_add(total, o);
            } else {
                total = o;
                init = true;
            }
        }
        if (!init)
            payload_zero(total);
        pending[n] = total;
    }
    __syncthreads();
}

// One breadth-first level [first, first + count) of one tree; levels run deepest first, so every
// internal child's total is complete (a kernel boundary orders the levels: no atomics, no fences).
template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, std::size_t first, std::size_t count, payload<Order> *__restrict__ pending,
    float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error
) {
    std::size_t const g = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    std::size_t const i = g >> 2;
    __shared__ payload<Order> share[k_block];
    __shared__ std::uint32_t share_have[k_block];
    moment_refit_slot<Order>(first + i, static_cast<int>(g & 3), i < count, nodes, node_count, tris, tri_count, pending,
                             records, slot_area, error, share, share_have);
}

// A run of small levels (at most k_fused_block nodes each) in one workgroup, a barrier between
// levels.
constexpr int k_fused_block = 512;
template <int Order>
__global__ __launch_bounds__(k_fused_block) void moment_levels_fused_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, level_run run, payload<Order> *__restrict__ pending, float4 *__restrict__ records,
    float *__restrict__ slot_area, std::uint32_t *error
) {
    __shared__ payload<Order> share[k_fused_block];
    __shared__ std::uint32_t share_have[k_fused_block];
    for (int k = 0; k < run.levels; ++k) {
        for (std::uint32_t base = 0; base < run.count[k]; base += k_fused_block / 4) {
            std::uint32_t const i = base + (threadIdx.x >> 2);
            moment_refit_slot<Order>(run.first[k] + i, static_cast<int>(threadIdx.x & 3), i < run.count[k], nodes,
                                     node_count, tris, tri_count, pending, records, slot_area, error, share, share_have);
        }
        __syncthreads();
    }
}

// Batched refits: one launch per depth over that depth's nodes in every mesh.  Thread g finds its
// segment (one mesh's node range at this depth) by binary search over the segments' prefixes.
__device__ inline std::uint32_t segment_of(refit_segment const *seg, std::uint32_t count, std::uint64_t g) {
    std::uint32_t lo = 0, hi = count;  // seg[lo].prefix <= g < seg[hi].prefix
    while (hi - lo > 1) {
        std::uint32_t const mid = lo + (hi - lo) / 2;
        if (seg[mid].prefix <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Thread t of a depth's quads (node t / 4, slot t % 4) over that depth's `total` nodes of every
// mesh; the whole block calls it.
template <int Order>
__device__ inline void moment_level_batch_item(
    moment_refit_mesh const *__restrict__ meshes, refit_segment const *__restrict__ seg, std::uint32_t seg_count,
    std::uint64_t total, std::uint64_t t, std::uint32_t *error, payload<Order> *share, std::uint32_t *share_have
) {
    std::uint64_t const g = t >> 2;  // node of this quad within the depth's nodes of every mesh
    bool active = g < total;
    moment_refit_mesh m{};
    std::size_t n = 0;
    if (active) {
        refit_segment const sg = seg[segment_of(seg, seg_count, g)];
        m = meshes[sg.mesh];
        n = static_cast<std::size_t>(sg.first_node) + static_cast<std::size_t>(g - sg.prefix);
        if (n >= m.node_count) {
            atomicOr(error, 1u);
            active = false;
        }
    }
    moment_refit_slot<Order>(n, static_cast<int>(t & 3), active, m.nodes, m.node_count, m.tris, m.tri_count,
                             static_cast<payload<Order> *>(m.pending), m.records, m.slot_area, error, share, share_have);
}

template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_batch_kernel(
    moment_refit_mesh const *__restrict__ meshes, refit_segment const *__restrict__ seg, std::uint32_t seg_count,
    std::uint64_t total, std::uint32_t *error
) {
    __shared__ payload<Order> share[k_block];
    __shared__ std::uint32_t share_have[k_block];
    moment_level_batch_item<Order>(meshes, seg, seg_count, total,
                                   static_cast<std::uint64_t>(blockIdx.x) * k_block + threadIdx.x, error, share,
                                   share_have);
}

// The same with the depth's segments and node count in device memory (written earlier in the
// stream by the segmented build), so the launch can be recorded before the tree exists: a fixed
// grid strides over the depth's quads.  The stride loop's bound is uniform over the block.
template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_batch_grid_kernel(
    moment_refit_mesh const *__restrict__ meshes, refit_segment const *__restrict__ segments,
    refit_level const *__restrict__ level, std::uint32_t *error
) {
    __shared__ payload<Order> share[k_block];
    __shared__ std::uint32_t share_have[k_block];
    refit_level const lv = *level;
    std::uint64_t const quads = 4 * lv.nodes;
    for (std::uint64_t base = static_cast<std::uint64_t>(blockIdx.x) * k_block; base < quads;
         base += static_cast<std::uint64_t>(gridDim.x) * k_block)
        moment_level_batch_item<Order>(meshes, segments + lv.first_segment, lv.segments, lv.nodes,
                                       base + threadIdx.x, error, share, share_have);
}

// Evaluation records from finished traversal records: one thread per (node, child slot) over
// nodes [0, node_count) of one tree (the child references and the zero padding are copied too).
template <int Order>
__device__ inline void eval_slot(float4 const *__restrict__ records, float4 *__restrict__ eval_records, std::size_t n,
                                 int s) {
    constexpr int RV = moment_record_vec4(Order);
    constexpr int EV = eval_record_vec4(Order);
    constexpr int B = eval_slot_floats(Order);
    constexpr int F = moment_field_count(Order);
    float const *src = reinterpret_cast<float const *>(records + n * RV);
    float *dst = reinterpret_cast<float *>(eval_records + n * EV);
    float c[F], e[F];
    for (int f = 0; f < F; ++f)
        c[f] = src[4 + f * 4 + s];
    taylor_eval_fields<Order>(c, e);
    // References and the acceptance fields (max_p_dist2, centre) keep their slot-interleaved place.
    dst[s] = src[s];
    for (int f = 0; f < 4; ++f)
        dst[4 + f * 4 + s] = e[f];
    float *blk = dst + 20 + s * B;
    for (int f = 4; f < F; ++f)
        blk[f - 4] = e[f];
    for (int k = F - 4; k < B; ++k)
        blk[k] = 0.0f;
}

template <int Order>
__global__ __launch_bounds__(256) void moment_eval_kernel(float4 const *__restrict__ records,
                                                          float4 *__restrict__ eval_records, std::size_t node_count) {
    std::size_t const g = static_cast<std::size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (g >= node_count * 4)
        return;
    eval_slot<Order>(records, eval_records, g >> 2, static_cast<int>(g & 3));
}

// The same over every mesh of a batched refit: thread g belongs to the last mesh whose slot prefix
// is <= g (binary search over `slot_prefix`, mesh_count + 1 entries).
template <int Order>
__global__ __launch_bounds__(256) void moment_eval_batch_kernel(moment_refit_mesh const *__restrict__ meshes,
                                                                std::uint64_t const *__restrict__ slot_prefix,
                                                                std::uint32_t mesh_count) {
    std::uint64_t const g = static_cast<std::uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (g >= slot_prefix[mesh_count])
        return;
    std::uint32_t lo = 0, hi = mesh_count;
    while (hi - lo > 1) {
        std::uint32_t const mid = lo + (hi - lo) / 2;
        if (slot_prefix[mid] <= g)
            lo = mid;
        else
            hi = mid;
    }
    moment_refit_mesh const m = meshes[lo];
    std::uint64_t const k = g - slot_prefix[lo];
    eval_slot<Order>(m.records, m.eval_records, static_cast<std::size_t>(k >> 2), static_cast<int>(k & 3));
}

// The same over a slot prefix in device memory (written by the segmented build), by a fixed grid.
template <int Order>
__global__ __launch_bounds__(256) void moment_eval_batch_grid_kernel(moment_refit_mesh const *__restrict__ meshes,
                                                                     std::uint64_t const *__restrict__ slot_prefix,
                                                                     std::uint32_t mesh_count) {
    std::uint64_t const total = slot_prefix[mesh_count];
    for (std::uint64_t g = static_cast<std::uint64_t>(blockIdx.x) * 256 + threadIdx.x; g < total;
         g += static_cast<std::uint64_t>(gridDim.x) * 256) {
        std::uint32_t lo = 0, hi = mesh_count;
        while (hi - lo > 1) {
            std::uint32_t const mid = lo + (hi - lo) / 2;
            if (slot_prefix[mid] <= g)
                lo = mid;
            else
                hi = mid;
        }
        moment_refit_mesh const m = meshes[lo];
        std::uint64_t const k = g - slot_prefix[lo];
        eval_slot<Order>(m.records, m.eval_records, static_cast<std::size_t>(k >> 2), static_cast<int>(k & 3));
    }
}

void launch_eval(gwn_moment_handle &m, std::size_t node_count, hipStream_t stream) {
    if (node_count == 0)
        return;
    unsigned const grid = grid_for(node_count * 4, 256);
    switch (m.order) {
    case 0: moment_eval_kernel<0><<<grid, 256, 0, stream>>>(m.records.ptr, m.eval_records.ptr, node_count); break;
    case 1: moment_eval_kernel<1><<<grid, 256, 0, stream>>>(m.records.ptr, m.eval_records.ptr, node_count); break;
    default: moment_eval_kernel<2><<<grid, 256, 0, stream>>>(m.records.ptr, m.eval_records.ptr, node_count); break;
    }
    GWN_LAUNCH_CHECK();
}

template <int Order>
void launch_levels(gwn_bvh_handle const &bvh, gwn_moment_handle &m, void *pending, std::uint32_t *error,
                   hipStream_t stream) {
    std::size_t const N = bvh.node_count;
    for (level_run const &r : plan_level_runs(bvh.level_start, k_fused_block)) {
        if (r.levels > 0)
            moment_levels_fused_kernel<Order><<<1, k_fused_block, 0, stream>>>(
                bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, r, static_cast<payload<Order> *>(pending),
                m.records.ptr, m.slot_area.ptr, error);
        else
            moment_level_kernel<Order><<<grid_for(4 * std::size_t(r.big_count), k_block), k_block, 0, stream>>>(
                bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, r.big_first, r.big_count,
                static_cast<payload<Order> *>(pending), m.records.ptr, m.slot_area.ptr, error);
        GWN_LAUNCH_CHECK();
    }
}

void check_levels(gwn_bvh_handle const &bvh) {
    auto const &ls = bvh.level_start;
    if (ls.size() < 2 || ls.front() != 0u || ls.back() != bvh.node_count)
        throw_internal("bvh.refit.moment: BVH level table does not cover its nodes.");
    for (std::size_t d = 1; d < ls.size(); ++d)
        if (ls[d] < ls[d - 1])
            throw_internal("bvh.refit.moment: BVH level table is not ascending.");
}

std::size_t payload_bytes(int order) {
    return order == 0 ? sizeof(payload<0>) : (order == 1 ? sizeof(payload<1>) : sizeof(payload<2>));
}

void enqueue_levels(gwn_bvh_handle const &bvh, gwn_moment_handle &m, void *pending, std::uint32_t *error,
                    hipStream_t stream) {
    switch (m.order) {
    case 0: launch_levels<0>(bvh, m, pending, error, stream); break;
    case 1: launch_levels<1>(bvh, m, pending, error, stream); break;
    case 2: launch_levels<2>(bvh, m, pending, error, stream); break;
    default: throw_invalid("Taylor order must be 0, 1, or 2.");
    }
}

} // namespace

void refit_moment_prepare(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream) {
    std::size_t const N = ref_internal(bvh.root.reference) ? bvh.node_count : 0;
    m.node_count = bvh.node_count;
    m.width = bvh.width;
    m.storage_generation = next_storage_generation();
    if (N == 0)
        return;
    check_levels(bvh);
    m.records.allocate(N * static_cast<std::size_t>(moment_record_vec4(m.order)), stream);
    m.eval_records.allocate(N * static_cast<std::size_t>(eval_record_vec4(m.order)), stream);
    m.slot_area.allocate(N * 4, stream);
    w.pending.allocate(N * payload_bytes(m.order), stream);
}

void refit_moment_batch(moment_refit_mesh const *meshes, refit_segment const *segments,
                        std::vector<refit_level> const &levels, std::uint64_t const *slot_prefix,
                        std::uint32_t mesh_count, int order, std::uint32_t *error, hipStream_t stream) {
    std::uint64_t total_nodes = 0;
    for (std::size_t d = levels.size(); d-- > 0;) {
        refit_level const &lv = levels[d];
        if (lv.nodes == 0)
            continue;
        unsigned const grid = grid_for(4 * lv.nodes, k_block);
        refit_segment const *seg = segments + lv.first_segment;
        switch (order) {
        case 0: moment_level_batch_kernel<0><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        case 1: moment_level_batch_kernel<1><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        case 2: moment_level_batch_kernel<2><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        default: throw_invalid("Taylor order must be 0, 1, or 2.");
        }
        GWN_LAUNCH_CHECK();
        total_nodes += lv.nodes;
    }
    if (total_nodes == 0 || mesh_count == 0)
        return;
    unsigned const grid = grid_for(4 * total_nodes, 256);
    switch (order) {
    case 0: moment_eval_batch_kernel<0><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    case 1: moment_eval_batch_kernel<1><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    default: moment_eval_batch_kernel<2><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    }
    GWN_LAUNCH_CHECK();
}

void refit_moment_batch_device(moment_refit_mesh const *meshes, refit_segment const *segments,
                               refit_level const *levels, int level_count, std::uint64_t const *slot_prefix,
                               std::uint32_t mesh_count, int order, std::uint32_t *error, unsigned grid,
                               hipStream_t stream) {
    if (order < 0 || order > 2)
        throw_invalid("Taylor order must be 0, 1, or 2.");
    if (mesh_count == 0 || grid == 0)
        return;
    for (int d = level_count; d-- > 0;) {
        switch (order) {
        case 0: moment_level_batch_grid_kernel<0><<<grid, k_block, 0, stream>>>(meshes, segments, levels + d, error); break;
        case 1: moment_level_batch_grid_kernel<1><<<grid, k_block, 0, stream>>>(meshes, segments, levels + d, error); break;
        default: moment_level_batch_grid_kernel<2><<<grid, k_block, 0, stream>>>(meshes, segments, levels + d, error); break;
        }
        GWN_LAUNCH_CHECK();
    }
    switch (order) {
    case 0: moment_eval_batch_grid_kernel<0><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    case 1: moment_eval_batch_grid_kernel<1><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    default: moment_eval_batch_grid_kernel<2><<<grid, 256, 0, stream>>>(meshes, slot_prefix, mesh_count); break;
    }
    GWN_LAUNCH_CHECK();
}

std::size_t moment_payload_bytes(int order) { return payload_bytes(order); }

void refit_moment(gwn_bvh_handle const &bvh, gwn_moment_handle &m, hipStream_t stream) {
    m.node_count = bvh.node_count;
    m.bvh_revision = bvh.revision;
    if (!ref_internal(bvh.root.reference) || bvh.node_count == 0)
        return; // a leaf root has no moment nodes (gwn_bvh.cuh:290-291)
    if (m.order < 0 || m.order > 2)
        throw_invalid("Taylor order must be 0, 1, or 2.");
    check_levels(bvh);
    std::size_t const N = bvh.node_count;
    m.records.allocate(N * static_cast<std::size_t>(moment_record_vec4(m.order)), stream);
    m.eval_records.allocate(N * static_cast<std::size_t>(eval_record_vec4(m.order)), stream);
    m.slot_area.allocate(N * 4, stream);
    device_buffer<std::uint8_t> pending(N * payload_bytes(m.order), stream);
    device_buffer<std::uint32_t> error(1, stream);
    fill_u32(error.ptr, 0u, 1, stream);
    enqueue_levels(bvh, m, pending.ptr, error.ptr, stream);
    launch_eval(m, N, stream);
    std::uint32_t host_error = 0;
    GWN_HIP(hipMemcpyAsync(&host_error, error.ptr, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.refit.moment: Bottom-up moment refit did not converge.");
}

} // namespace gwn
