// gwn/gwn_accessor.hpp - typed accessors of libgwn_hip.so's objects, with the reference's names.
//
// An accessor is a trivially copyable, non-owning snapshot of an object's device storage
// (include/gwn/gwn_geometry.cuh:19-58, gwn_bvh.cuh:209-329, gwn_boundary.cuh:28-54 of the
// reference).  The C ABI fills the plain structs of gwn_hip.h; these class templates derive from
// them to carry the reference's template parameters, so host code (gwn/gwn.hpp, any C++ compiler)
// can return them and device code (gwn/gwn_device.hpp, hipcc) can take them by value in a kernel.
// Any upload, build, refit, clear or destroy of the object invalidates a snapshot.
#pragma once

#include <cstdint>

#include "../gwn_hip.h"

#if defined(__HIPCC__)
#define GWN_HOST_DEVICE __host__ __device__
#else
#define GWN_HOST_DEVICE
#endif

namespace gwn {

// Default traversal stack capacity of batches and point queries (include/gwn/gwn_query.cuh:45).
inline constexpr int k_gwn_default_traversal_stack_capacity = 64;

template <class Real = float, class Index = std::uint32_t> struct gwn_geometry_accessor : ::gwn_geometry_accessor {
    static_assert(sizeof(Real) == 4 && Real(0.5) != Real(0), "libgwn_hip provides Real = float");
    static_assert(sizeof(Index) == 4 && Index(-1) > Index(0), "libgwn_hip provides Index = uint32_t");
    [[nodiscard]] GWN_HOST_DEVICE std::uint64_t vertex_count() const noexcept { return this->::gwn_geometry_accessor::vertex_count; }
    [[nodiscard]] GWN_HOST_DEVICE std::uint64_t triangle_count() const noexcept {
        return this->::gwn_geometry_accessor::triangle_count;
    }
    [[nodiscard]] GWN_HOST_DEVICE bool is_valid() const noexcept {
        return x != nullptr && y != nullptr && z != nullptr && i0 != nullptr && i1 != nullptr && i2 != nullptr;
    }
};

// gwn_bvh_accessor<Width, Real, Index>: nodes are stored four-wide for every width (slots past a
// node's children are empty), records are (v0, v1, v2) float triples in leaf order.
template <int Width, class Real = float, class Index = std::uint32_t> struct gwn_bvh_accessor : ::gwn_bvh_accessor {
    static_assert(Width >= 2 && Width <= 4, "libgwn_hip builds BVHs of width 2, 3 or 4");
    static_assert(sizeof(Real) == 4 && Real(0.5) != Real(0), "libgwn_hip provides Real = float");
    static_assert(sizeof(Index) == 4 && Index(-1) > Index(0), "libgwn_hip provides Index = uint32_t");
    [[nodiscard]] GWN_HOST_DEVICE bool has_internal_root() const noexcept {
        return (root_reference >> 47 & 1u) != 0 && (root_reference >> 48) == 0;
    }
    [[nodiscard]] GWN_HOST_DEVICE bool has_leaf_root() const noexcept {
        return (root_reference >> 47 & 1u) != 0 && (root_reference >> 48) != 0;
    }
    // Queryable: a current revision and a valid root (include/gwn/gwn_bvh.cuh:230-258).
    [[nodiscard]] GWN_HOST_DEVICE bool is_valid() const noexcept {
        return revision != 0 && triangle_count != 0 && (has_internal_root() || has_leaf_root());
    }
};

template <class Real = float, class Index = std::uint32_t> using gwn_bvh4_accessor = gwn_bvh_accessor<4, Real, Index>;

template <int Width, int Order, class Real = float, class Index = std::uint32_t>
struct gwn_bvh_moment_accessor : ::gwn_moment_accessor {
    static_assert(Order >= 0 && Order <= 2, "Taylor order must be 0, 1, or 2");
    // Formed from `bvh`'s current revision (include/gwn/gwn_bvh.cuh:286-293).
    [[nodiscard]] GWN_HOST_DEVICE bool is_valid_for(gwn_bvh_accessor<Width, Real, Index> const &bvh) const noexcept {
        return order == Order && bvh_revision != 0 && bvh_revision == bvh.revision &&
               (records != nullptr || bvh.has_leaf_root());
    }
};

template <class Index = std::uint32_t> struct gwn_boundary_chain_accessor : ::gwn_boundary_accessor {
    static_assert(sizeof(Index) == 4 && Index(-1) > Index(0), "libgwn_hip provides Index = uint32_t");
    [[nodiscard]] GWN_HOST_DEVICE bool is_valid() const noexcept { return is_built != 0; }
};

template <class Real> struct gwn_vec3 {
    Real x, y, z;
};

enum class gwn_ray_first_hit_status : std::uint8_t { k_miss = 0, k_hit = 1, k_overflow = 2 };

// gwn_ray_first_hit_result<Real, Index> (include/gwn/gwn_query.cuh:63-95): 32 bytes, the layout
// of the batch's records (t = -1 and primitive_id = 0xFFFFFFFF on a miss).
template <class Real = float, class Index = std::uint32_t> struct gwn_ray_first_hit_result {
    Real t{Real(-1)};
    Real u{Real(0)};
    Real v{Real(0)};
    Real geometric_normal_x{Real(0)};
    Real geometric_normal_y{Real(0)};
    Real geometric_normal_z{Real(0)};
    Index primitive_id{Index(0xffffffffu)};
    gwn_ray_first_hit_status status{gwn_ray_first_hit_status::k_miss};
    [[nodiscard]] GWN_HOST_DEVICE constexpr bool hit() const noexcept { return status == gwn_ray_first_hit_status::k_hit; }

private:
    std::uint8_t storage_padding_[3]{};
};

} // namespace gwn
