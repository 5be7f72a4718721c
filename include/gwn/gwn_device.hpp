// gwn/gwn_device.hpp - device point queries for user kernels (HIP, gfx950), with the reference's
// names: gwn_unsigned_distance, gwn_ray_first_hit, gwn_winding_number_exact,
// gwn_winding_number_taylor, gwn_winding_gradient_taylor, gwn_winding_number_antipodal and
// gwn_winding_gradient_antipodal (include/gwn/gwn_query.cuh:114-127, 189-204, 296-302, 354-378,
// 459-483, 566-584, 656-682 of the reference).
//
// Each function evaluates one query in the calling thread, with no block-level cooperation, from
// typed accessors of libgwn_hip.so's objects (gwn/gwn_accessor.hpp; pass them to a kernel by
// value).  The arithmetic is the library's own per-lane query code (gwn/detail/gwn_device_*.hpp),
// so a result equals the corresponding batch's result for the same query bit for bit when this
// header is compiled with the library's flags (hipcc --offload-arch=gfx950
// -fhip-fp32-correctly-rounded-divide-sqrt); the contraction mode is set inside every function.
//
// Overflow contract (include/gwn/detail/gwn_query_common_impl.cuh:10-29): a traversal that would
// exceed StackCapacity calls overflow_callback() once and returns NaN (distance, Taylor winding,
// antipodal winding), an all-NaN vector (Taylor gradient), or the best hit found so far with status
// k_overflow (first hit).  The default callback traps, as the reference's does; pass a functor
// with `__device__ void operator()() const` to record the event instead.  The batch entry points
// validate the capacity up front and never overflow.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gwn_accessor.hpp"
#include "detail/gwn_device_antipodal.hpp"
#include "detail/gwn_device_distance.hpp"
#include "detail/gwn_device_layout.hpp"
#include "detail/gwn_device_ray.hpp"
#include "detail/gwn_device_winding.hpp"

namespace gwn {

namespace detail {
// The reference's default overflow callback (gwn_query_common_impl.cuh:10-13): stop the thread.
struct gwn_traversal_overflow_trap_callback {
    __device__ void operator()() const noexcept { __builtin_trap(); }
};
} // namespace detail

namespace detail {

template <int Width, class Real, class Index>
__device__ inline tri_record const *records_of(gwn_bvh_accessor<Width, Real, Index> const &bvh) {
    return reinterpret_cast<tri_record const *>(bvh.triangles);
}

template <int Width, class Real, class Index>
__device__ inline bvh_node4 const *nodes_of(gwn_bvh_accessor<Width, Real, Index> const &bvh) {
    return static_cast<bvh_node4 const *>(bvh.nodes);
}

} // namespace detail

// Unsigned distance to the mesh (include/gwn/gwn_query.cuh:114-127): culling_band is the initial
// bound and the result when no triangle is closer (negative values clamp to zero).
template <int Width, class Real, class Index, int StackCapacity = k_gwn_default_traversal_stack_capacity,
          class OverflowCallback = detail::gwn_traversal_overflow_trap_callback>
[[nodiscard]] __device__ inline Real gwn_unsigned_distance(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, Real const qx, Real const qy, Real const qz,
    Real const culling_band = Real(__builtin_huge_valf()), OverflowCallback const &overflow_callback = {}
) noexcept {
#pragma clang fp contract(off)
    static_assert(StackCapacity > 0, "Traversal stack capacity must be positive.");
    namespace dd = detail::distance;
    dd::distance_scene const sc{detail::nodes_of(bvh), bvh.node_count, detail::records_of(bvh), bvh.triangle_count,
                                bvh.root_reference};
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    int const st = dd::distance_traverse<StackCapacity>(sc, qx, qy, qz, best, nullptr);
    if (st == dd::k_dist_overflow) {
        overflow_callback();
        return __builtin_nanf("");
    }
    return st == dd::k_dist_zero ? 0.0f : sqrtf(best);
}

// Nearest hit on [t_min, t_max] (include/gwn/gwn_query.cuh:189-204): the packed near-first
// traversal when the BVH's packed stack bound fits StackCapacity, else the canonical one -- the
// batch's rule, so equal-t ties resolve as the batch resolves them.
template <int Width, class Real, class Index, int StackCapacity = k_gwn_default_traversal_stack_capacity,
          class OverflowCallback = detail::gwn_traversal_overflow_trap_callback>
[[nodiscard]] __device__ inline gwn_ray_first_hit_result<Real, Index> gwn_ray_first_hit(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, Real const ray_ox, Real const ray_oy, Real const ray_oz,
    Real const ray_dx, Real const ray_dy, Real const ray_dz, Real const t_min = Real(0),
    Real const t_max = Real(__builtin_huge_valf()), OverflowCallback const &overflow_callback = {}
) noexcept {
    static_assert(StackCapacity > 0, "Traversal stack capacity must be positive.");
    namespace dr = detail::ray;
    dr::ray_scene const sc{detail::nodes_of(bvh), bvh.node_count, detail::records_of(bvh), bvh.primitive_indices,
                           bvh.triangle_count};
    bool const packed = bvh.has_internal_root() && bvh.node_count < (std::uint64_t(1) << 29) &&
                        bvh.packed_stack_bound <= static_cast<std::uint32_t>(StackCapacity);
    dr::ray_hit const h =
        packed ? dr::ray_first_hit_lane<true, StackCapacity>(sc, bvh.root_reference, ray_ox, ray_oy, ray_oz, ray_dx,
                                                             ray_dy, ray_dz, t_min, t_max, nullptr)
               : dr::ray_first_hit_lane<false, StackCapacity>(sc, bvh.root_reference, ray_ox, ray_oy, ray_oz, ray_dx,
                                                              ray_dy, ray_dz, t_min, t_max, nullptr);
    if (h.status == 2)
        overflow_callback();
    gwn_ray_first_hit_result<Real, Index> r{};
    r.t = h.t;
    r.u = h.u;
    r.v = h.v;
    r.geometric_normal_x = h.nx;
    r.geometric_normal_y = h.ny;
    r.geometric_normal_z = h.nz;
    r.primitive_id = h.primitive_id;
    r.status = static_cast<gwn_ray_first_hit_status>(h.status);
    return r;
}

// Exact winding number over every triangle record in leaf order (include/gwn/gwn_query.cuh:296-302):
// the plain batch's float sum, term by term.
template <int Width, class Real, class Index>
[[nodiscard]] __device__ inline Real gwn_winding_number_exact(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, Real const qx, Real const qy, Real const qz
) noexcept {
#pragma clang fp contract(on)
    namespace dw = detail::winding;
    tri_record const *tris = detail::records_of(bvh);
    float sum = 0.0f;
    for (std::uint64_t k = 0; k < bvh.triangle_count; ++k)
        sum += dw::solid_angle_plain(tris[k], qx, qy, qz);
    return sum / (4.0f * dw::k_pi_f);
}

namespace detail {

template <int Order, bool Grad, int StackCapacity, int Width, class Real, class Index, class OverflowCallback>
__device__ inline winding::taylor_sums<GWN_PRECISION_PLAIN, Grad> taylor_point(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, gwn_bvh_moment_accessor<Width, Order, Real, Index> const &moment,
    float x, float y, float z, float accuracy_scale, OverflowCallback const &overflow_callback, bool &overflow
) {
#pragma clang fp contract(on)
    winding::taylor_sums<GWN_PRECISION_PLAIN, Grad> acc;
    overflow = false;
    if (bvh.has_leaf_root()) {
        // A one-leaf tree: every triangle exactly, as the batch's leaf-root kernel does.
        std::uint64_t const off = bvh.root_reference & ((std::uint64_t(1) << 47) - 1);
        std::uint32_t const cnt = static_cast<std::uint32_t>(bvh.root_reference >> 48);
        for (std::uint32_t k = 0; k < cnt; ++k)
            acc.leaf(records_of(bvh)[off + k], x, y, z);
        return acc;
    }
    float const beta2 = accuracy_scale * accuracy_scale;
    std::uint32_t n_tests = 0, n_accept = 0, n_leaf = 0;
    // The value alone reads the evaluation records, the gradient the traversal records, as the
    // batches do.
    overflow = winding::taylor_traverse<Order, GWN_PRECISION_PLAIN, Grad, StackCapacity>(
        static_cast<float4 const *>(Grad ? moment.records : moment.eval_records), records_of(bvh),
        static_cast<std::uint32_t>(bvh.root_reference & ((std::uint64_t(1) << 47) - 1)), x, y, z, beta2, acc,
        n_tests, n_accept, n_leaf);
    if (overflow)
        overflow_callback();
    return acc;
}

} // namespace detail

// Taylor-approximated winding number (include/gwn/gwn_query.cuh:354-378): the per-lane traversal
// of the batch with the reference's acceptance test |q - P|^2 > r^2 accuracy_scale^2.
template <int Order, int Width, class Real, class Index, int StackCapacity = k_gwn_default_traversal_stack_capacity,
          class OverflowCallback = detail::gwn_traversal_overflow_trap_callback>
[[nodiscard]] __device__ inline Real gwn_winding_number_taylor(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, gwn_bvh_moment_accessor<Width, Order, Real, Index> const &moment,
    Real const qx, Real const qy, Real const qz, Real const accuracy_scale = Real(2),
    OverflowCallback const &overflow_callback = {}
) noexcept {
#pragma clang fp contract(on)
    static_assert(StackCapacity > 0, "Traversal stack capacity must be positive.");
    bool overflow = false;
    auto const acc = detail::taylor_point<Order, false, StackCapacity>(bvh, moment, qx, qy, qz, accuracy_scale,
                                                                       overflow_callback, overflow);
    if (overflow)
        return __builtin_nanf("");
    return acc.omega / (4.0f * detail::winding::k_pi_f);
}

// Gradient of the Taylor winding number (include/gwn/gwn_query.cuh:656-682).
template <int Order, int Width, class Real, class Index, int StackCapacity = k_gwn_default_traversal_stack_capacity,
          class OverflowCallback = detail::gwn_traversal_overflow_trap_callback>
[[nodiscard]] __device__ inline gwn_vec3<Real> gwn_winding_gradient_taylor(
    gwn_bvh_accessor<Width, Real, Index> const &bvh, gwn_bvh_moment_accessor<Width, Order, Real, Index> const &moment,
    Real const qx, Real const qy, Real const qz, Real const accuracy_scale = Real(2),
    OverflowCallback const &overflow_callback = {}
) noexcept {
#pragma clang fp contract(on)
    static_assert(StackCapacity > 0, "Traversal stack capacity must be positive.");
    bool overflow = false;
    auto const acc = detail::taylor_point<Order, true, StackCapacity>(bvh, moment, qx, qy, qz, accuracy_scale,
                                                                      overflow_callback, overflow);
    if (overflow) {
        float const nan = __builtin_nanf("");
        return gwn_vec3<Real>{nan, nan, nan};
    }
    return gwn_vec3<Real>{acc.gf[0] * detail::winding::k_inv_4pi_f, acc.gf[1] * detail::winding::k_inv_4pi_f,
                          acc.gf[2] * detail::winding::k_inv_4pi_f};
}

namespace detail {

template <class Real, class Index>
__device__ inline antipodal::boundary_view boundary_of(gwn_geometry_accessor<Real, Index> const &geometry,
                                                       gwn_boundary_chain_accessor<Index> const &chain) {
    return antipodal::boundary_view{geometry.x, geometry.y, geometry.z, chain.start, chain.end, chain.multiplicity,
                                    chain.edge_count};
}

} // namespace detail

// Antipodal winding number (include/gwn/gwn_query.cuh:459-483): crossings of the +Z ray (then +X,
// then +Y when an axis is singular) plus the boundary chain's spherical areas over 4 pi; NaN when
// every axis is singular.
template <int Width, class Real, class Index, int StackCapacity = k_gwn_default_traversal_stack_capacity,
          class OverflowCallback = detail::gwn_traversal_overflow_trap_callback>
[[nodiscard]] __device__ inline Real gwn_winding_number_antipodal(
    gwn_geometry_accessor<Real, Index> const &geometry, gwn_bvh_accessor<Width, Real, Index> const &bvh,
    gwn_boundary_chain_accessor<Index> const &boundary_chain, Real const qx, Real const qy, Real const qz,
    OverflowCallback const &overflow_callback = {}
) noexcept {
    static_assert(StackCapacity > 0, "Traversal stack capacity must be positive.");
    namespace da = detail::antipodal;
    da::crossing_scene const sc{detail::nodes_of(bvh), bvh.node_count, detail::records_of(bvh), bvh.triangle_count,
                                bvh.root_reference};
    bool overflowed = false;
    float const w = da::antipodal_point<GWN_PRECISION_PLAIN, StackCapacity>(
        sc, detail::boundary_of(geometry, boundary_chain), da::v3{qx, qy, qz}, overflowed);
    if (overflowed)
        overflow_callback();
    return w;
}

// Gradient of the antipodal boundary term (include/gwn/gwn_query.cuh:566-584): (NaN, NaN, NaN)
// when every axis is singular.
template <class Real, class Index>
[[nodiscard]] __device__ inline gwn_vec3<Real> gwn_winding_gradient_antipodal(
    gwn_geometry_accessor<Real, Index> const &geometry, gwn_boundary_chain_accessor<Index> const &boundary_chain,
    Real const qx, Real const qy, Real const qz
) noexcept {
    float g[3];
    detail::antipodal::antipodal_gradient_point<GWN_PRECISION_PLAIN>(
        detail::boundary_of(geometry, boundary_chain), detail::antipodal::v3{qx, qy, qz}, g);
    return gwn_vec3<Real>{g[0], g[1], g[2]};
}

} // namespace gwn
