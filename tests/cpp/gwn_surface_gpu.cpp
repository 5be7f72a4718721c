// GPU workflow through the reference-named C++ surface (include/gwn/gwn.hpp), built as a small
// shared library that the GPU tests load: upload, build (width 2, 3 or 4; LBVH or H-PLOC), order-2
// moment refit, exact and Taylor winding batches, then an in-place position update, bounds and
// moment refits and the Taylor batch again, all on one non-blocking stream of the caller's
// process, with the objects' stream bindings checked along the way (the reference's workflow,
// tests/integration_gwn_dynamic.cu:259-330).  Results are copied to host arrays for comparison
// with the oracle.
#include <gwn/gwn.hpp>

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::string g_message;

template <class T> struct device_array {
    T *ptr = nullptr;
    explicit device_array(std::size_t n) {
        if (n && hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)) != hipSuccess)
            ptr = nullptr;
    }
    ~device_array() {
        if (ptr)
            (void)hipFree(ptr);
    }
    device_array(device_array const &) = delete;
    device_array &operator=(device_array const &) = delete;
};

#define SURFACE_CHECK(cond, what)                                                                  \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            g_message = what;                                                                      \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

#define SURFACE_STATUS(expr)                                                                       \
    do {                                                                                           \
        gwn::gwn_status const st_ = (expr);                                                        \
        if (!st_.is_ok()) {                                                                        \
            g_message = std::string(#expr) + ": " + st_.message();                                 \
            return 2;                                                                              \
        }                                                                                          \
    } while (0)

template <int Width, class Real, class Index>
int run(Real const *vx, Real const *vy, Real const *vz, std::uint64_t nv, Index const *i0, Index const *i1,
        Index const *i2, std::uint64_t nt, Real const *moved_x, Real const *moved_y, Real const *moved_z,
        Real const *qx, Real const *qy, Real const *qz, std::uint64_t nq, int method, std::uint32_t radius,
        Real beta, Real *exact, Real *taylor, Real *refit_taylor, std::uint64_t *info) {
    hipStream_t stream = nullptr;
    SURFACE_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess, "stream create");
    struct stream_guard {
        hipStream_t s;
        ~stream_guard() { (void)hipStreamDestroy(s); }
    } guard{stream};

    gwn::gwn_geometry_object<Real, Index> geometry;
    SURFACE_STATUS(gwn::gwn_upload_geometry(
        geometry, gwn::gwn_host_span<Real const>(vx, nv), gwn::gwn_host_span<Real const>(vy, nv),
        gwn::gwn_host_span<Real const>(vz, nv), gwn::gwn_host_span<Index const>(i0, nt),
        gwn::gwn_host_span<Index const>(i1, nt), gwn::gwn_host_span<Index const>(i2, nt), stream));
    SURFACE_CHECK(geometry.stream() == stream, "geometry bound to the upload stream");
    gwn::gwn_bvh_object<Width, Real, Index> bvh;
    gwn::gwn_bvh_build_options options{};
    options.method = method == 0 ? gwn::gwn_bvh_build_method::k_hploc : gwn::gwn_bvh_build_method::k_lbvh;
    options.hploc_search_radius = radius;
    SURFACE_STATUS(gwn::gwn_build_bvh(geometry, bvh, options, stream));
    SURFACE_CHECK(bvh.stream() == stream && bvh.has_data(), "BVH built and bound");
    SURFACE_CHECK(bvh.info().width == static_cast<std::uint32_t>(Width), "BVH width");
    gwn::gwn_bvh_moment_object<Width, 2, Real, Index> moment;
    SURFACE_STATUS(gwn::gwn_refit_bvh_moment<2>(bvh, moment, stream));
    SURFACE_CHECK(moment.is_valid_for(bvh), "moment aligned with the BVH");

    device_array<Real> dq[3] = {device_array<Real>(nq), device_array<Real>(nq), device_array<Real>(nq)};
    device_array<Real> dout(nq);
    Real const *hq[3] = {qx, qy, qz};
    for (int k = 0; k < 3; ++k)
        SURFACE_CHECK(hipMemcpyAsync(dq[k].ptr, hq[k], nq * sizeof(Real), hipMemcpyHostToDevice, stream) ==
                          hipSuccess, "query copy");
    gwn::gwn_device_span<Real const> const sx(dq[0].ptr, nq), sy(dq[1].ptr, nq), sz(dq[2].ptr, nq);
    gwn::gwn_device_span<Real> const so(dout.ptr, nq);
    constexpr gwn::gwn_query_batch_config config{256, 128};
    SURFACE_STATUS((gwn::gwn_compute_winding_number_exact_batch<config>(bvh, sx, sy, sz, so, stream)));
    SURFACE_CHECK(hipMemcpyAsync(exact, dout.ptr, nq * sizeof(Real), hipMemcpyDeviceToHost, stream) == hipSuccess,
                  "exact copy");
    SURFACE_STATUS((gwn::gwn_compute_winding_number_taylor_batch<2, config>(bvh, moment, sx, sy, sz, so, beta,
                                                                           stream)));
    SURFACE_CHECK(hipMemcpyAsync(taylor, dout.ptr, nq * sizeof(Real), hipMemcpyDeviceToHost, stream) == hipSuccess,
                  "Taylor copy");

    // Dynamic workflow: update positions, refit bounds (new revision: the moment is stale until
    // refit), refit moments, query again.
    std::uint64_t const revision = bvh.info().revision;
    SURFACE_STATUS(gwn::gwn_update_geometry(geometry, gwn::gwn_host_span<Real const>(moved_x, nv),
                                            gwn::gwn_host_span<Real const>(moved_y, nv),
                                            gwn::gwn_host_span<Real const>(moved_z, nv), stream));
    SURFACE_STATUS(gwn::gwn_refit_bvh(geometry, bvh, stream));
    SURFACE_CHECK(bvh.info().revision != revision && !moment.is_valid_for(bvh), "refit issues a new revision");
    gwn::gwn_status const stale =
        gwn::gwn_compute_winding_number_taylor_batch<2, config>(bvh, moment, sx, sy, sz, so, beta, stream);
    SURFACE_CHECK(stale.error() == gwn::gwn_error::invalid_argument, "stale moment refused");
    SURFACE_STATUS(gwn::gwn_refit_bvh_moment<2>(bvh, moment, stream));
    SURFACE_STATUS((gwn::gwn_compute_winding_number_taylor_batch<2, config>(bvh, moment, sx, sy, sz, so, beta,
                                                                           stream)));
    SURFACE_CHECK(hipMemcpyAsync(refit_taylor, dout.ptr, nq * sizeof(Real), hipMemcpyDeviceToHost, stream) ==
                      hipSuccess, "refit Taylor copy");
    SURFACE_CHECK(hipStreamSynchronize(stream) == hipSuccess, "stream synchronize");
    gwn_bvh_info const i = bvh.info();
    info[0] = i.node_count;
    info[1] = i.internal_stack_bound;
    info[2] = i.packed_stack_bound;
    info[3] = i.width;
    // Release on the bound stream, explicitly.
    moment.clear();
    bvh.clear();
    geometry.clear();
    SURFACE_CHECK(!bvh.has_data() && !moment.has_data() && !geometry.has_data(), "clear releases");
    SURFACE_CHECK(hipStreamSynchronize(stream) == hipSuccess, "stream synchronize after clear");
    return 0;
}

} // namespace

extern "C" int gwn_surface_workflow(
    float const *vx, float const *vy, float const *vz, std::uint64_t nv, std::uint32_t const *i0,
    std::uint32_t const *i1, std::uint32_t const *i2, std::uint64_t nt, float const *moved_x, float const *moved_y,
    float const *moved_z, float const *qx, float const *qy, float const *qz, std::uint64_t nq, int width,
    int method, float *exact, float *taylor, float *refit_taylor, std::uint64_t *info) {
    g_message.clear();
    switch (width) {
    case 2: return run<2>(vx, vy, vz, nv, i0, i1, i2, nt, moved_x, moved_y, moved_z, qx, qy, qz, nq, method, 8u,
                          2.0f, exact, taylor, refit_taylor, info);
    case 3: return run<3>(vx, vy, vz, nv, i0, i1, i2, nt, moved_x, moved_y, moved_z, qx, qy, qz, nq, method, 8u,
                          2.0f, exact, taylor, refit_taylor, info);
    case 4: return run<4>(vx, vy, vz, nv, i0, i1, i2, nt, moved_x, moved_y, moved_z, qx, qy, qz, nq, method, 8u,
                          2.0f, exact, taylor, refit_taylor, info);
    default: g_message = "width must be 2, 3 or 4"; return 3;
    }
}

// The same workflow on gwn_geometry_object<double, std::uint64_t> and its width-4 BVH and order-2
// moments (the reference's FP64 builder workflow, tests/integration_gwn_builder.cu:345-435).
extern "C" int gwn_surface_workflow_f64_u64(
    double const *vx, double const *vy, double const *vz, std::uint64_t nv, std::uint64_t const *i0,
    std::uint64_t const *i1, std::uint64_t const *i2, std::uint64_t nt, double const *moved_x,
    double const *moved_y, double const *moved_z, double const *qx, double const *qy, double const *qz,
    std::uint64_t nq, int method, std::uint32_t radius, double beta, double *exact, double *taylor,
    double *refit_taylor, std::uint64_t *info) {
    g_message.clear();
    return run<4>(vx, vy, vz, nv, i0, i1, i2, nt, moved_x, moved_y, moved_z, qx, qy, qz, nq, method, radius, beta,
                  exact, taylor, refit_taylor, info);
}

extern "C" char const *gwn_surface_message() { return g_message.c_str(); }
