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

// One node of a level pass: the payload of each child slot (leaf payloads straight from the
// triangle records, internal ones from the child's total written by the previous, deeper level),
// the node's recentred coefficients and slot areas, and the node's total in slot order from the
// first valid slot (the reference's combine_node, gwn_bvh_refit_moment_impl.cuh:17-99).  The whole
// record is written, empty slots and padding as zeros, so the output needs no clearing pass.
template <int Order>
__device__ inline void moment_refit_node(
    std::size_t n, bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, payload<Order> *__restrict__ pending, float4 *__restrict__ records,
    float *__restrict__ slot_area, std::uint32_t *error
) {
    constexpr int RV = moment_record_vec4(Order);
    constexpr int F = moment_field_count(Order);
    float4 *rec = records + n * RV;
    float *fields = reinterpret_cast<float *>(rec + 1);
    std::uint32_t cref[4] = {k_invalid, k_invalid, k_invalid, k_invalid};
    payload<Order> total;
    payload_zero(total);
    bool init = false;
    for (int s = 0; s < 4; ++s) {
        bvh_child const ch = nodes[n].children[s];
        std::uint64_t const ref = ch.reference;
        if (!ref_valid(ref)) {
            for (int f = 0; f < F; ++f)
                fields[f * 4 + s] = 0.0f;
            slot_area[n * 4 + s] = 0.0f;
            continue;
        }
        payload<Order> child;
        if (ref_internal(ref)) {
            std::uint64_t const c = ref_offset(ref);
            if (c >= node_count || c <= n) {
                atomicOr(error, 1u);
                return;
            }
            child = pending[c];
            cref[s] = static_cast<std::uint32_t>(c);
        } else {
            // A leaf payload starts from zero and adds each primitive in order
            // (gwn_bvh_refit_common.cuh:452-468).
            std::uint64_t const off = ref_offset(ref);
            std::uint32_t const cnt = ref_count(ref);
            if (off + cnt > tri_count || cnt != 1u) {
                atomicOr(error, 1u);
                return;
            }
            payload_zero(child);
            payload<Order> tri;
            triangle_payload<Order>(tris[off], tri);
            payload_add(child, tri);
            cref[s] = k_cref_leaf | static_cast<std::uint32_t>(off);
        }
        write_fields<Order>(fields, s, ch.bounds, child);
        slot_area[n * 4 + s] = child.v[0];
        if (init)
            payload_add(total, child);
        else
            total = child;
        init = true;
    }
    for (int k = 4 + F * 4; k < RV * 4; ++k)
        reinterpret_cast<float *>(rec)[k] = 0.0f;
    rec[0] = make_float4(__uint_as_float(cref[0]), __uint_as_float(cref[1]), __uint_as_float(cref[2]),
                         __uint_as_float(cref[3]));
    pending[n] = total;
}

// One breadth-first level [first, first + count) of one tree; levels run deepest first, so every
// internal child's total is complete (a kernel boundary orders the levels: no atomics, no fences).
template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, std::size_t first, std::size_t count, payload<Order> *__restrict__ pending,
    float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= count)
        return;
    moment_refit_node<Order>(first + i, nodes, node_count, tris, tri_count, pending, records, slot_area, error);
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

template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_batch_kernel(
    moment_refit_mesh const *__restrict__ meshes, refit_segment const *__restrict__ seg, std::uint32_t seg_count,
    std::uint64_t total, std::uint32_t *error
) {
    std::uint64_t const g = static_cast<std::uint64_t>(blockIdx.x) * k_block + threadIdx.x;
    if (g >= total)
        return;
    refit_segment const sg = seg[segment_of(seg, seg_count, g)];
    moment_refit_mesh const m = meshes[sg.mesh];
    std::size_t const n = static_cast<std::size_t>(sg.first_node) + static_cast<std::size_t>(g - sg.prefix);
    if (n >= m.node_count) {
        atomicOr(error, 1u);
        return;
    }
    moment_refit_node<Order>(n, m.nodes, m.node_count, m.tris, m.tri_count, static_cast<payload<Order> *>(m.pending),
                             m.records, m.slot_area, error);
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
    auto const &ls = bvh.level_start;
    for (std::size_t d = ls.size() - 1; d-- > 0;) {
        std::size_t const first = ls[d], count = ls[d + 1] - ls[d];
        if (count == 0)
            continue;
        moment_level_kernel<Order><<<grid_for(count, k_block), k_block, 0, stream>>>(
            bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, first, count,
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
        unsigned const grid = grid_for(lv.nodes, k_block);
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
