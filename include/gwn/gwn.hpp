// gwn/gwn.hpp - C++ surface of libgwn_hip.so with the reference's names (kririae/smallgwn).
//
// Header-only.  Each template below keeps the name, template parameters and argument order of
// the reference function it replaces and forwards to one extern "C" entry point of
// include/gwn_hip.h, so code written against the reference's public API compiles against this
// header.  Objects take Real = float or double, Index = std::uint32_t or std::uint64_t and
// Width = 2, 3 or 4.  Every query batch takes Real = float; Real = double objects are uploaded,
// built, refit, moment-refit and queried by the exact and Taylor winding batches (the reference's
// FP64 workflow); the other double batches and the device accessors of double objects are
// rejected at compile time.  Needs C++20 (class-type template parameters, as the reference uses).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>

#include "../gwn_hip.h"
#include "gwn_accessor.hpp"

namespace gwn {

// ---- status (include/gwn/gwn_utils.cuh:213-283) ---------------------------------------------

enum class gwn_error {
    success = GWN_SUCCESS,
    invalid_argument = GWN_ERROR_INVALID_ARGUMENT,
    cuda_runtime_error = GWN_ERROR_HIP_RUNTIME, // the HIP runtime error on this platform
    internal_error = GWN_ERROR_INTERNAL,
};

class gwn_status {
public:
    gwn_status() = default;
    gwn_status(gwn_error e, std::string message) : error_(e), message_(std::move(message)) {}
    [[nodiscard]] static gwn_status ok() { return {}; }
    [[nodiscard]] bool is_ok() const noexcept { return error_ == gwn_error::success; }
    [[nodiscard]] gwn_error error() const noexcept { return error_; }
    [[nodiscard]] std::string const &message() const noexcept { return message_; }

private:
    gwn_error error_ = gwn_error::success;
    std::string message_;
};

namespace detail {
template <class Real> inline constexpr bool gwn_is_real_v =
    (sizeof(Real) == 4 || sizeof(Real) == 8) && Real(0.5) != Real(0);
template <class Real> inline constexpr bool gwn_is_double_v = gwn_is_real_v<Real> && sizeof(Real) == 8;
template <class Index> inline constexpr bool gwn_is_index_v =
    (sizeof(Index) == 4 || sizeof(Index) == 8) && Index(-1) > Index(0);
template <class Real> inline constexpr bool gwn_float_only_v = sizeof(Real) == 4;

[[nodiscard]] inline gwn_status gwn_from_code(int code) {
    if (code == GWN_SUCCESS)
        return gwn_status::ok();
    char const *msg = gwn_last_error_message();
    return gwn_status(static_cast<gwn_error>(code), msg ? msg : "");
}
} // namespace detail

// The reference's default stream is the legacy default stream (gwn_utils.cuh:45).
[[nodiscard]] inline hipStream_t gwn_default_stream() noexcept { return nullptr; }

// ---- spans (include/gwn/gwn_utils.cuh:84-171): explicit host vs device memory ----------------

template <class T> class gwn_host_span final {
public:
    constexpr gwn_host_span() noexcept = default;
    constexpr gwn_host_span(T *data, std::size_t size) noexcept : data_(data), size_(size) {}
    [[nodiscard]] constexpr T *data() const noexcept { return data_; }
    [[nodiscard]] constexpr std::size_t size() const noexcept { return size_; }

private:
    T *data_ = nullptr;
    std::size_t size_ = 0;
};

template <class T> class gwn_device_span final {
public:
    constexpr gwn_device_span() noexcept = default;
    constexpr gwn_device_span(T *data, std::size_t size) noexcept : data_(data), size_(size) {}
    template <class U>
    constexpr gwn_device_span(gwn_device_span<U> other) noexcept : data_(other.data()), size_(other.size()) {}
    [[nodiscard]] constexpr T *data() const noexcept { return data_; }
    [[nodiscard]] constexpr std::size_t size() const noexcept { return size_; }

private:
    T *data_ = nullptr;
    std::size_t size_ = 0;
};

// ---- objects: move-only owners of device state (gwn_geometry.cuh:98-149, gwn_bvh.cuh:331-429) --

namespace detail {
// Move-only owner of one library handle with the reference's stream mixin
// (include/gwn/gwn_utils.cuh:59-80): stream() is the bound stream, set_stream() rebinds without
// synchronizing, clear() releases the storage on the bound stream, and swap exchanges storage and
// binding together.  The binding lives in the library object, so it follows the operations that
// replace or write the storage exactly as the C ABI documents.
template <class Handle, int (*Destroy)(Handle), int (*GetStream)(Handle, hipStream_t *),
          int (*SetStream)(Handle, hipStream_t), int (*Clear)(Handle)>
class gwn_handle_owner {
public:
    gwn_handle_owner() = default;
    gwn_handle_owner(gwn_handle_owner const &) = delete;
    gwn_handle_owner &operator=(gwn_handle_owner const &) = delete;
    gwn_handle_owner(gwn_handle_owner &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    gwn_handle_owner &operator=(gwn_handle_owner &&o) noexcept {
        std::swap(h_, o.h_);
        return *this;
    }
    ~gwn_handle_owner() {
        if (h_ != nullptr)
            Destroy(h_);
    }
    [[nodiscard]] Handle get() const noexcept { return h_; }
    [[nodiscard]] hipStream_t stream() const noexcept {
        hipStream_t s = nullptr;
        GetStream(h_, &s);
        return s;
    }
    void set_stream(hipStream_t stream) noexcept { SetStream(h_, stream); }
    void clear() noexcept { Clear(h_); }
    friend void swap(gwn_handle_owner &a, gwn_handle_owner &b) noexcept { std::swap(a.h_, b.h_); }

protected:
    Handle h_ = nullptr;
};
} // namespace detail

template <class Real, class Index>
class gwn_geometry_object final
    : public detail::gwn_handle_owner<gwn_geometry_t, gwn_geometry_destroy, gwn_geometry_stream,
                                      gwn_geometry_set_stream, gwn_geometry_clear> {
    static_assert(detail::gwn_is_real_v<Real>, "libgwn_hip provides Real = float or double");
    static_assert(detail::gwn_is_index_v<Index>, "libgwn_hip provides Index = uint32_t or uint64_t");

public:
    gwn_geometry_object() { gwn_geometry_create(&h_); }
    [[nodiscard]] std::size_t vertex_count() const noexcept {
        std::uint64_t v = 0, t = 0;
        gwn_geometry_counts(h_, &v, &t);
        return v;
    }
    [[nodiscard]] std::size_t triangle_count() const noexcept {
        std::uint64_t v = 0, t = 0;
        gwn_geometry_counts(h_, &v, &t);
        return t;
    }
    [[nodiscard]] gwn_geometry_accessor<Real, Index> accessor() const noexcept {
        gwn_geometry_accessor<Real, Index> a{};
        gwn_geometry_get_accessor(h_, &a);
        return a;
    }
    [[nodiscard]] bool has_data() const noexcept { return vertex_count() != 0 || triangle_count() != 0; }
};

template <int Width, class Real, class Index>
class gwn_bvh_object final
    : public detail::gwn_handle_owner<gwn_bvh_t, gwn_bvh_destroy, gwn_bvh_stream, gwn_bvh_set_stream,
                                      gwn_bvh_clear> {
    static_assert(Width >= 2 && Width <= 4, "libgwn_hip builds BVHs of width 2, 3 or 4");
    static_assert(detail::gwn_is_real_v<Real>, "libgwn_hip provides Real = float or double");
    static_assert(detail::gwn_is_index_v<Index>, "libgwn_hip provides Index = uint32_t or uint64_t");

public:
    gwn_bvh_object() { gwn_bvh_create(&h_); }
    [[nodiscard]] gwn_bvh_info info() const noexcept {
        gwn_bvh_info i{};
        gwn_bvh_get_info(h_, &i);
        return i;
    }
    [[nodiscard]] gwn_bvh_accessor<Width, Real, Index> accessor() const noexcept {
        gwn_bvh_accessor<Width, Real, Index> a{};
        gwn_bvh_get_accessor(h_, &a);
        return a;
    }
    [[nodiscard]] bool has_data() const noexcept { return info().triangle_count != 0; }
};

template <int Width, int Order, class Real, class Index>
class gwn_bvh_moment_object final
    : public detail::gwn_handle_owner<gwn_moment_t, gwn_moment_destroy, gwn_moment_stream, gwn_moment_set_stream,
                                      gwn_moment_clear> {
    static_assert(Width >= 2 && Width <= 4, "libgwn_hip builds BVHs of width 2, 3 or 4");
    static_assert(Order >= 0 && Order <= 2, "Taylor order must be 0, 1, or 2");

public:
    gwn_bvh_moment_object() { gwn_moment_create(&h_, Order); }
    [[nodiscard]] gwn_bvh_moment_accessor<Width, Order, Real, Index> accessor() const noexcept {
        gwn_bvh_moment_accessor<Width, Order, Real, Index> a{};
        gwn_moment_get_accessor(h_, &a);
        return a;
    }
    // Valid for `bvh` when formed from its current revision (include/gwn/gwn_bvh.cuh:286-293): a
    // queryable BVH, the same revision, and moment nodes unless the root is a leaf.
    [[nodiscard]] bool is_valid_for(gwn_bvh_object<Width, Real, Index> const &bvh) const noexcept {
        std::uint64_t nodes = 0, revision = 0;
        gwn_moment_info(h_, &nodes, &revision);
        gwn_bvh_info const i = bvh.info();
        bool const leaf_root = (i.root_reference >> 47 & 1u) != 0 && (i.root_reference >> 48) != 0;
        return revision != 0 && revision == i.revision && i.triangle_count != 0 && (nodes != 0 || leaf_root);
    }
    [[nodiscard]] bool has_data() const noexcept {
        std::uint64_t nodes = 0;
        gwn_moment_info(h_, &nodes, nullptr);
        return nodes != 0;
    }
};

// gwn_boundary_chain_object<Index> (include/gwn/gwn_boundary.cuh:56-125).
template <class Index = std::uint32_t>
class gwn_boundary_chain_object final
    : public detail::gwn_handle_owner<gwn_boundary_t, gwn_boundary_destroy, gwn_boundary_stream,
                                      gwn_boundary_set_stream, gwn_boundary_clear> {
    static_assert(sizeof(Index) == 4 && Index(-1) > Index(0), "libgwn_hip provides Index = uint32_t");

public:
    gwn_boundary_chain_object() { gwn_boundary_create(&h_); }
    [[nodiscard]] std::size_t edge_count() const noexcept {
        std::uint64_t e = 0;
        gwn_boundary_info(h_, &e, nullptr, nullptr, nullptr);
        return e;
    }
    [[nodiscard]] bool is_built() const noexcept {
        std::int32_t b = 0;
        gwn_boundary_info(h_, nullptr, nullptr, nullptr, &b);
        return b != 0;
    }
    [[nodiscard]] bool has_data() const noexcept { return is_built(); }
    [[nodiscard]] gwn_boundary_chain_accessor<Index> accessor() const noexcept {
        gwn_boundary_chain_accessor<Index> a{};
        gwn_boundary_get_accessor(h_, &a);
        return a;
    }
};

template <class Real = float, class Index = std::uint32_t> using gwn_bvh4_object = gwn_bvh_object<4, Real, Index>;
template <int Order, class Real = float, class Index = std::uint32_t>
using gwn_bvh4_moment_object = gwn_bvh_moment_object<4, Order, Real, Index>;

// ---- geometry (include/gwn/gwn_geometry.cuh:60-90) ------------------------------------------

template <class Real, class Index>
[[nodiscard]] gwn_status gwn_upload_geometry(
    gwn_geometry_object<Real, Index> &object, gwn_host_span<Real const> x, gwn_host_span<Real const> y,
    gwn_host_span<Real const> z, gwn_host_span<Index const> i0, gwn_host_span<Index const> i1,
    gwn_host_span<Index const> i2, hipStream_t stream = gwn_default_stream()
) noexcept {
    if (y.size() != x.size() || z.size() != x.size() || i1.size() != i0.size() || i2.size() != i0.size())
        return gwn_status(gwn_error::invalid_argument, "Geometry SoA spans must have identical lengths.");
    int code;
    if constexpr (detail::gwn_is_double_v<Real> && sizeof(Index) == 8)
        code = gwn_upload_geometry_f64_u64(object.get(), x.data(), y.data(), z.data(), x.size(),
                                           reinterpret_cast<std::uint64_t const *>(i0.data()),
                                           reinterpret_cast<std::uint64_t const *>(i1.data()),
                                           reinterpret_cast<std::uint64_t const *>(i2.data()), i0.size(), stream);
    else if constexpr (detail::gwn_is_double_v<Real>)
        code = gwn_upload_geometry_f64(object.get(), x.data(), y.data(), z.data(), x.size(),
                                       reinterpret_cast<std::uint32_t const *>(i0.data()),
                                       reinterpret_cast<std::uint32_t const *>(i1.data()),
                                       reinterpret_cast<std::uint32_t const *>(i2.data()), i0.size(), stream);
    else if constexpr (sizeof(Index) == 8)
        code = gwn_upload_geometry_f32_u64(object.get(), x.data(), y.data(), z.data(), x.size(),
                                           reinterpret_cast<std::uint64_t const *>(i0.data()),
                                           reinterpret_cast<std::uint64_t const *>(i1.data()),
                                           reinterpret_cast<std::uint64_t const *>(i2.data()), i0.size(), stream);
    else
        code = gwn_upload_geometry_f32(object.get(), x.data(), y.data(), z.data(), x.size(),
                                       reinterpret_cast<std::uint32_t const *>(i0.data()),
                                       reinterpret_cast<std::uint32_t const *>(i1.data()),
                                       reinterpret_cast<std::uint32_t const *>(i2.data()), i0.size(), stream);
    return detail::gwn_from_code(code);
}

template <class Real, class Index>
[[nodiscard]] gwn_status gwn_update_geometry(
    gwn_geometry_object<Real, Index> &object, gwn_host_span<Real const> x, gwn_host_span<Real const> y,
    gwn_host_span<Real const> z, hipStream_t stream = gwn_default_stream()
) noexcept {
    if (y.size() != x.size() || z.size() != x.size())
        return gwn_status(gwn_error::invalid_argument, "Vertex SoA spans must have identical lengths.");
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(gwn_update_geometry_f64(object.get(), x.data(), y.data(), z.data(), x.size(), stream));
    else
        return detail::gwn_from_code(gwn_update_geometry_f32(object.get(), x.data(), y.data(), z.data(), x.size(), stream));
}

// ---- build and refit (include/gwn/gwn_bvh_build.cuh:16-39, gwn_bvh_refit.cuh:25-51) ----------

enum class gwn_bvh_build_method : std::uint8_t { k_hploc, k_lbvh };

struct gwn_bvh_build_options {
    gwn_bvh_build_method method = gwn_bvh_build_method::k_hploc;
    std::uint32_t hploc_search_radius = 8;
};

template <int Width, class Real, class Index, class MortonCode = std::uint64_t>
[[nodiscard]] gwn_status gwn_build_bvh(
    gwn_geometry_object<Real, Index> const &geometry, gwn_bvh_object<Width, Real, Index> &bvh,
    gwn_bvh_build_options const options = {}, hipStream_t stream = gwn_default_stream()
) noexcept {
    int const method = options.method == gwn_bvh_build_method::k_hploc ? GWN_BVH_BUILD_HPLOC : GWN_BVH_BUILD_LBVH;
    // MortonCode: codes are always 63-bit (uint64_t, the reference default); a uint32_t request
    // builds the same tree order the library supports and is accepted for source compatibility.
    static_assert(sizeof(MortonCode) == 4 || sizeof(MortonCode) == 8, "MortonCode must be uint32_t or uint64_t");
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(
            gwn_build_bvh_f64(geometry.get(), bvh.get(), Width, method, options.hploc_search_radius, stream)
        );
    else
        return detail::gwn_from_code(
            gwn_build_bvh_f32(geometry.get(), bvh.get(), Width, method, options.hploc_search_radius, stream)
        );
}

template <int Width, class Real, class Index>
[[nodiscard]] gwn_status gwn_refit_bvh(
    gwn_geometry_object<Real, Index> const &geometry, gwn_bvh_object<Width, Real, Index> &bvh,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(gwn_refit_bvh_f64(geometry.get(), bvh.get(), stream));
    else
        return detail::gwn_from_code(gwn_refit_bvh_f32(geometry.get(), bvh.get(), stream));
}

template <int Order, int Width, class Real, class Index>
[[nodiscard]] gwn_status gwn_refit_bvh_moment(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_bvh_moment_object<Width, Order, Real, Index> &moment,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(gwn_refit_bvh_moment_f64(bvh.get(), moment.get(), stream));
    else
        return detail::gwn_from_code(gwn_refit_bvh_moment_f32(bvh.get(), moment.get(), stream));
}

// ---- query batches (include/gwn/gwn_query.cuh:41-61, 296-445) --------------------------------

inline constexpr int k_gwn_default_query_batch_block_size = 256;

struct gwn_query_batch_config {
    int block_size = k_gwn_default_query_batch_block_size;
    int stack_capacity = k_gwn_default_traversal_stack_capacity;
};

namespace detail {
template <gwn_query_batch_config Config> [[nodiscard]] constexpr ::gwn_query_config gwn_c_config() {
    return ::gwn_query_config{Config.block_size, Config.stack_capacity, -1, 0};
}
} // namespace detail

template <gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real, class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024)
[[nodiscard]] gwn_status gwn_compute_winding_number_exact_batch(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_device_span<Real const> query_x,
    gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z, gwn_device_span<Real> output,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    if (query_y.size() != query_x.size() || query_z.size() != query_x.size() || output.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Query SoA and output spans must have identical lengths.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(gwn_compute_winding_number_exact_batch_f64(
            bvh.get(), query_x.data(), query_y.data(), query_z.data(), output.data(), output.size(), &c, stream
        ));
    else
        return detail::gwn_from_code(gwn_compute_winding_number_exact_batch_f32(
            bvh.get(), query_x.data(), query_y.data(), query_z.data(), output.data(), output.size(), &c, stream
        ));
}

template <int Order, gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real,
          class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024 && Config.stack_capacity > 0)
[[nodiscard]] gwn_status gwn_compute_winding_number_taylor_batch(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_bvh_moment_object<Width, Order, Real, Index> const &moment,
    gwn_device_span<Real const> query_x, gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z,
    gwn_device_span<Real> output, Real const accuracy_scale = Real(2), hipStream_t stream = gwn_default_stream()
) noexcept {
    if (query_y.size() != query_x.size() || query_z.size() != query_x.size() || output.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Query SoA and output spans must have identical lengths.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    if constexpr (detail::gwn_is_double_v<Real>)
        return detail::gwn_from_code(gwn_compute_winding_number_taylor_batch_f64(
            bvh.get(), moment.get(), query_x.data(), query_y.data(), query_z.data(), output.data(), output.size(),
            accuracy_scale, &c, stream
        ));
    else
        return detail::gwn_from_code(gwn_compute_winding_number_taylor_batch_f32(
            bvh.get(), moment.get(), query_x.data(), query_y.data(), query_z.data(), output.data(), output.size(),
            accuracy_scale, &c, stream
        ));
}

template <int Order, gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real,
          class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024 && Config.stack_capacity > 0)
[[nodiscard]] gwn_status gwn_compute_winding_gradient_taylor_batch(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_bvh_moment_object<Width, Order, Real, Index> const &moment,
    gwn_device_span<Real const> query_x, gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z,
    gwn_device_span<Real> output_x, gwn_device_span<Real> output_y, gwn_device_span<Real> output_z,
    Real const accuracy_scale = Real(2), hipStream_t stream = gwn_default_stream()
) noexcept {
    static_assert(detail::gwn_float_only_v<Real>, "libgwn_hip provides this batch for Real = float");
    if (query_y.size() != query_x.size() || query_z.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Query SoA spans must have identical lengths.");
    if (output_x.size() != query_x.size() || output_y.size() != query_x.size() || output_z.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Output span sizes must match query count.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    return detail::gwn_from_code(gwn_compute_winding_gradient_taylor_batch_f32(
        bvh.get(), moment.get(), query_x.data(), query_y.data(), query_z.data(), output_x.data(), output_y.data(),
        output_z.data(), output_x.size(), accuracy_scale, &c, stream
    ));
}

template <gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real, class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024 && Config.stack_capacity > 0)
[[nodiscard]] gwn_status gwn_compute_unsigned_distance_batch(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_device_span<Real const> query_x,
    gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z, gwn_device_span<Real> output,
    Real const culling_band = Real(__builtin_huge_valf()), hipStream_t stream = gwn_default_stream()
) noexcept {
    static_assert(detail::gwn_float_only_v<Real>, "libgwn_hip provides this batch for Real = float");
    if (query_y.size() != query_x.size() || query_z.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Query SoA spans must have identical lengths.");
    if (output.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Output span size must match query count.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    return detail::gwn_from_code(gwn_compute_unsigned_distance_batch_f32(
        bvh.get(), query_x.data(), query_y.data(), query_z.data(), output.data(), output.size(), culling_band, &c,
        stream
    ));
}

static_assert(sizeof(gwn_ray_first_hit_result<>) == sizeof(::gwn_ray_first_hit_result));

template <gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real, class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024 && Config.stack_capacity > 0)
[[nodiscard]] gwn_status gwn_compute_ray_first_hit_batch(
    gwn_bvh_object<Width, Real, Index> const &bvh, gwn_device_span<Real const> ray_origin_x,
    gwn_device_span<Real const> ray_origin_y, gwn_device_span<Real const> ray_origin_z,
    gwn_device_span<Real const> ray_dir_x, gwn_device_span<Real const> ray_dir_y,
    gwn_device_span<Real const> ray_dir_z, gwn_device_span<gwn_ray_first_hit_result<Real, Index>> output,
    Real const t_min = Real(0), Real const t_max = Real(__builtin_huge_valf()),
    hipStream_t stream = gwn_default_stream()
) noexcept {
    static_assert(detail::gwn_float_only_v<Real>, "libgwn_hip provides this batch for Real = float");
    std::size_t const n = ray_origin_x.size();
    if (ray_origin_y.size() != n || ray_origin_z.size() != n || ray_dir_x.size() != n || ray_dir_y.size() != n ||
        ray_dir_z.size() != n || output.size() != n)
        return gwn_status(gwn_error::invalid_argument, "ray first-hit: mismatched span sizes");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    return detail::gwn_from_code(gwn_compute_ray_first_hit_batch_f32(
        bvh.get(), ray_origin_x.data(), ray_origin_y.data(), ray_origin_z.data(), ray_dir_x.data(), ray_dir_y.data(),
        ray_dir_z.data(), reinterpret_cast<::gwn_ray_first_hit_result *>(output.data()), n, t_min, t_max, &c, stream
    ));
}

// ---- boundary chain and antipodal batches (include/gwn/gwn_boundary.cuh:103-125,
//      include/gwn/gwn_query.cuh:492-640) ----------------------------------------------------

template <class Real, class Index>
[[nodiscard]] gwn_status gwn_build_boundary_chain(
    gwn_geometry_object<Real, Index> const &geometry, gwn_boundary_chain_object<Index> &chain,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    return detail::gwn_from_code(::gwn_build_boundary_chain(geometry.get(), chain.get(), stream));
}

template <gwn_query_batch_config Config = gwn_query_batch_config{}, int Width, class Real, class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024 && Config.stack_capacity > 0)
[[nodiscard]] gwn_status gwn_compute_winding_number_antipodal_batch(
    gwn_geometry_object<Real, Index> const &geometry, gwn_bvh_object<Width, Real, Index> const &bvh,
    gwn_boundary_chain_object<Index> const &chain, gwn_device_span<Real const> query_x,
    gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z, gwn_device_span<Real> output,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    static_assert(detail::gwn_float_only_v<Real>, "libgwn_hip provides this batch for Real = float");
    if (query_y.size() != query_x.size() || query_z.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Query SoA spans require identical lengths.");
    if (output.size() != query_x.size())
        return gwn_status(gwn_error::invalid_argument, "Output span size differs from query count.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    return detail::gwn_from_code(gwn_compute_winding_number_antipodal_batch_f32(
        geometry.get(), bvh.get(), chain.get(), query_x.data(), query_y.data(), query_z.data(), output.data(),
        output.size(), &c, stream
    ));
}

template <gwn_query_batch_config Config = gwn_query_batch_config{}, class Real, class Index = std::uint32_t>
    requires(Config.block_size > 0 && Config.block_size <= 1024)
[[nodiscard]] gwn_status gwn_compute_winding_gradient_antipodal_batch(
    gwn_geometry_object<Real, Index> const &geometry, gwn_boundary_chain_object<Index> const &chain,
    gwn_device_span<Real const> query_x, gwn_device_span<Real const> query_y, gwn_device_span<Real const> query_z,
    gwn_device_span<Real> output_x, gwn_device_span<Real> output_y, gwn_device_span<Real> output_z,
    hipStream_t stream = gwn_default_stream()
) noexcept {
    static_assert(detail::gwn_float_only_v<Real>, "libgwn_hip provides this batch for Real = float");
    std::size_t const n = query_x.size();
    if (query_y.size() != n || query_z.size() != n)
        return gwn_status(gwn_error::invalid_argument, "Query SoA spans require identical lengths.");
    if (output_x.size() != n || output_y.size() != n || output_z.size() != n)
        return gwn_status(gwn_error::invalid_argument, "Output span sizes differ from query count.");
    constexpr ::gwn_query_config c = detail::gwn_c_config<Config>();
    return detail::gwn_from_code(gwn_compute_winding_gradient_antipodal_batch_f32(
        geometry.get(), chain.get(), query_x.data(), query_y.data(), query_z.data(), output_x.data(),
        output_y.data(), output_z.data(), n, &c, stream
    ));
}

} // namespace gwn
