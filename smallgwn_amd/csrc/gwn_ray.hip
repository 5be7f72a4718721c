// gwn_ray.hip - nearest ray-triangle hit batch on gfx950.
//
// Reference: include/gwn/detail/gwn_query_ray_impl.cuh:20-595, include/gwn/gwn_query.cuh:182-290.
//
// One ray per lane.  Two traversals, chosen per batch exactly as the reference's launcher does:
//   packed    (internal root, fewer than 2^29 nodes, packed_stack_bound within the stack): leaf
//             and internal children share one near-ordered sequence of 32-bit payloads (leaf =
//             bit 31 | parent << 2 | slot); the nearest child is entered, the others pushed;
//   canonical (otherwise): leaves of a node are intersected in near order before its internal
//             children are queued; only internal node ids occupy the stack.
// A hit replaces the current best only with a strictly smaller t, so ties resolve by traversal
// order -- the same order as the reference's traversal of the same tree.
//
// Arithmetic is unfused with correctly rounded division, so every record equals the CPU
// oracle's replay bit for bit.
#include "gwn_core.hpp"
#include "../../include/gwn/detail/gwn_device_ray.hpp"

#include <cfloat>

#pragma clang fp contract(off)

namespace gwn {
namespace {

using namespace ::gwn::detail::ray;
constexpr int k_block = 256;
// With Count, only the work counters (slab tests, triangle tests) are written, two per ray.
template <bool Packed, int Cap, bool Count = false>
__global__ __launch_bounds__(k_block) void ray_kernel(
    ray_scene sc, std::uint64_t root_ref, float const *__restrict__ ox, float const *__restrict__ oy,
    float const *__restrict__ oz, float const *__restrict__ dx, float const *__restrict__ dy,
    float const *__restrict__ dz, ray_hit *__restrict__ out, std::size_t n, float t_min, float t_max,
    std::uint32_t *__restrict__ counters = nullptr
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    traversal_work w;
    ray_hit const res = ray_first_hit_lane<Packed, Cap>(sc, root_ref, ox[i], oy[i], oz[i], dx[i], dy[i], dz[i], t_min,
                                                        t_max, Count ? &w : nullptr);
    if constexpr (Count) {
        counters[2 * i] = w.boxes;
        counters[2 * i + 1] = w.tris;
    } else {
        out[i] = res;
    }
}

template <bool Packed, int Cap>
void launch_ray(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, void *out, std::size_t n, float t_min, float t_max,
    std::uint32_t *counters, hipStream_t stream
) {
    ray_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.primitive_indices.ptr, bvh.triangle_count};
    if (counters != nullptr)
        ray_kernel<Packed, Cap, true><<<grid_for(n, k_block), k_block, 0, stream>>>(
            sc, bvh.root.reference, ox, oy, oz, dx, dy, dz, nullptr, n, t_min, t_max, counters);
    else
        ray_kernel<Packed, Cap><<<grid_for(n, k_block), k_block, 0, stream>>>(
            sc, bvh.root.reference, ox, oy, oz, dx, dy, dz, static_cast<ray_hit *>(out), n, t_min, t_max);
    GWN_LAUNCH_CHECK();
}

template <bool Packed>
void launch_ray_cap(
    gwn_bvh_handle const &bvh, int need, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, void *out, std::size_t n, float t_min, float t_max, std::uint32_t *counters,
    hipStream_t stream
) {
    if (need <= 32)
        launch_ray<Packed, 32>(bvh, ox, oy, oz, dx, dy, dz, out, n, t_min, t_max, counters, stream);
    else if (need <= 64)
        launch_ray<Packed, 64>(bvh, ox, oy, oz, dx, dy, dz, out, n, t_min, t_max, counters, stream);
    else if (need <= 128)
        launch_ray<Packed, 128>(bvh, ox, oy, oz, dx, dy, dz, out, n, t_min, t_max, counters, stream);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
}

void dispatch_ray(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, void *out, std::size_t n, float t_min, float t_max,
    query_launch const &cfg, std::uint32_t *counters, hipStream_t stream
) {
    if (n == 0)
        return;
    // gwn_query.cuh:276-289: packed payloads hold a 29-bit parent offset.
    bool const packed = ref_internal(bvh.root.reference) && bvh.node_count < (std::size_t(1) << 29) &&
                        static_cast<int>(bvh.packed_stack_bound) <= cfg.stack_capacity;
    if (packed)
        launch_ray_cap<true>(bvh, static_cast<int>(bvh.packed_stack_bound), ox, oy, oz, dx, dy, dz, out, n, t_min,
                             t_max, counters, stream);
    else
        launch_ray_cap<false>(bvh, static_cast<int>(bvh.internal_stack_bound), ox, oy, oz, dx, dy, dz, out, n,
                              t_min, t_max, counters, stream);
}

} // namespace

void ray_first_hit(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, void *out, std::size_t n, float t_min, float t_max,
    query_launch const &cfg, hipStream_t stream
) {
    dispatch_ray(bvh, ox, oy, oz, dx, dy, dz, out, n, t_min, t_max, cfg, nullptr, stream);
}

void ray_first_hit_work(
    gwn_bvh_handle const &bvh, float const *ox, float const *oy, float const *oz, float const *dx,
    float const *dy, float const *dz, std::size_t n, float t_min, float t_max, query_launch const &cfg,
    std::uint32_t *counters, hipStream_t stream
) {
    dispatch_ray(bvh, ox, oy, oz, dx, dy, dz, nullptr, n, t_min, t_max, cfg, counters, stream);
}

} // namespace gwn
