// Test kernels for the device point API (include/gwn/gwn_device.hpp): one query per thread through
// gwn_unsigned_distance, gwn_ray_first_hit, gwn_winding_number_exact, gwn_winding_number_taylor,
// gwn_winding_gradient_taylor, gwn_winding_number_antipodal and gwn_winding_gradient_antipodal,
// with a callback that records stack overflow per query (the reference's overflow_probe,
// tests/unit_gwn_query_distance_ray.cu:86-110).  Built by tests/cpp/build.py with the library's
// flags; the GPU tests compare every result with the corresponding batch's.
#include <gwn/gwn.hpp>
#include <gwn/gwn_device.hpp>

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

struct overflow_probe {
    int *flags = nullptr;
    std::uint64_t index = 0;
    __device__ void operator()() const noexcept {
        if (flags != nullptr)
            flags[index] = 1;
    }
};

using bvh_acc = gwn::gwn_bvh_accessor<4, float, std::uint32_t>;
using geo_acc = gwn::gwn_geometry_accessor<float, std::uint32_t>;
using chain_acc = gwn::gwn_boundary_chain_accessor<std::uint32_t>;

struct queries {
    float const *x, *y, *z, *dx, *dy, *dz;
    std::uint64_t n;
};

__global__ void exact_kernel(bvh_acc bvh, queries q, float *out) {
    std::uint64_t const i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < q.n)
        out[i] = gwn::gwn_winding_number_exact(bvh, q.x[i], q.y[i], q.z[i]);
}

template <int Order, int Cap>
__global__ void taylor_kernel(bvh_acc bvh, gwn::gwn_bvh_moment_accessor<4, Order, float, std::uint32_t> mom, queries q,
                              float beta, float *out, float *gx, float *gy, float *gz, int *flags) {
    std::uint64_t const i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= q.n)
        return;
    out[i] = gwn::gwn_winding_number_taylor<Order, 4, float, std::uint32_t, Cap>(bvh, mom, q.x[i], q.y[i], q.z[i], beta,
                                                                                 overflow_probe{flags, i});
    gwn::gwn_vec3<float> const g = gwn::gwn_winding_gradient_taylor<Order, 4, float, std::uint32_t, Cap>(
        bvh, mom, q.x[i], q.y[i], q.z[i], beta, overflow_probe{});
    gx[i] = g.x;
    gy[i] = g.y;
    gz[i] = g.z;
}

template <int Cap>
__global__ void distance_kernel(bvh_acc bvh, queries q, float band, float *out, int *flags) {
    std::uint64_t const i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < q.n)
        out[i] = gwn::gwn_unsigned_distance<4, float, std::uint32_t, Cap>(bvh, q.x[i], q.y[i], q.z[i], band,
                                                                          overflow_probe{flags, i});
}

template <int Cap>
__global__ void ray_kernel(bvh_acc bvh, queries q, float t_min, float t_max,
                           gwn::gwn_ray_first_hit_result<float, std::uint32_t> *out, int *flags) {
    std::uint64_t const i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < q.n)
        out[i] = gwn::gwn_ray_first_hit<4, float, std::uint32_t, Cap>(bvh, q.x[i], q.y[i], q.z[i], q.dx[i], q.dy[i],
                                                                      q.dz[i], t_min, t_max, overflow_probe{flags, i});
}

template <int Cap>
__global__ void antipodal_kernel(geo_acc geo, bvh_acc bvh, chain_acc chain, queries q, float *out, float *gx,
                                 float *gy, float *gz, int *flags) {
    std::uint64_t const i = std::uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= q.n)
        return;
    out[i] = gwn::gwn_winding_number_antipodal<4, float, std::uint32_t, Cap>(geo, bvh, chain, q.x[i], q.y[i], q.z[i],
                                                                             overflow_probe{flags, i});
    gwn::gwn_vec3<float> const g = gwn::gwn_winding_gradient_antipodal(geo, chain, q.x[i], q.y[i], q.z[i]);
    gx[i] = g.x;
    gy[i] = g.y;
    gz[i] = g.z;
}

unsigned grid(std::uint64_t n) { return static_cast<unsigned>((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

template <int Order, int Cap>
void launch_taylor(gwn_bvh_t b, gwn_moment_t m, queries q, float beta, float *out, float *gx, float *gy, float *gz,
                   int *flags, hipStream_t s) {
    bvh_acc ba{};
    gwn_bvh_get_accessor(b, &ba);
    gwn::gwn_bvh_moment_accessor<4, Order, float, std::uint32_t> ma{};
    gwn_moment_get_accessor(m, &ma);
    taylor_kernel<Order, Cap><<<grid(q.n), 256, 0, s>>>(ba, ma, q, beta, out, gx, gy, gz, flags);
}

} // namespace

// op: 0 exact, 1 Taylor + gradient (order, beta), 2 distance (band), 3 first hit (t_min = a,
// t_max = b), 4 antipodal winding + gradient.  cap: the compiled stack capacity, 64 or 1 (1 makes
// deep traversals overflow, for the callback contract).  Returns a hipError_t.
extern "C" int gwn_point_run(int op, int order, int cap, gwn_geometry_t g, gwn_bvh_t b, gwn_moment_t m,
                             gwn_boundary_t c, float const *qx, float const *qy, float const *qz, float const *dx,
                             float const *dy, float const *dz, std::uint64_t n, float a, float bb, void *out,
                             float *gx, float *gy, float *gz, int *flags, hipStream_t stream) {
    if (n == 0)
        return hipSuccess;
    queries const q{qx, qy, qz, dx, dy, dz, n};
    bvh_acc ba{};
    if (b != nullptr)
        gwn_bvh_get_accessor(b, &ba);
    float *fo = static_cast<float *>(out);
    bool const deep = cap != 1;
    switch (op) {
    case 0: exact_kernel<<<grid(n), 256, 0, stream>>>(ba, q, fo); break;
    case 1:
        if (order == 0)
            deep ? launch_taylor<0, 64>(b, m, q, a, fo, gx, gy, gz, flags, stream)
                 : launch_taylor<0, 1>(b, m, q, a, fo, gx, gy, gz, flags, stream);
        else if (order == 1)
            deep ? launch_taylor<1, 64>(b, m, q, a, fo, gx, gy, gz, flags, stream)
                 : launch_taylor<1, 1>(b, m, q, a, fo, gx, gy, gz, flags, stream);
        else
            deep ? launch_taylor<2, 64>(b, m, q, a, fo, gx, gy, gz, flags, stream)
                 : launch_taylor<2, 1>(b, m, q, a, fo, gx, gy, gz, flags, stream);
        break;
    case 2:
        if (deep)
            distance_kernel<64><<<grid(n), 256, 0, stream>>>(ba, q, a, fo, flags);
        else
            distance_kernel<1><<<grid(n), 256, 0, stream>>>(ba, q, a, fo, flags);
        break;
    case 3: {
        auto *ro = static_cast<gwn::gwn_ray_first_hit_result<float, std::uint32_t> *>(out);
        if (deep)
            ray_kernel<64><<<grid(n), 256, 0, stream>>>(ba, q, a, bb, ro, flags);
        else
            ray_kernel<1><<<grid(n), 256, 0, stream>>>(ba, q, a, bb, ro, flags);
        break;
    }
    case 4: {
        geo_acc ga{};
        gwn_geometry_get_accessor(g, &ga);
        chain_acc ca{};
        gwn_boundary_get_accessor(c, &ca);
        if (deep)
            antipodal_kernel<64><<<grid(n), 256, 0, stream>>>(ga, ba, ca, q, fo, gx, gy, gz, flags);
        else
            antipodal_kernel<1><<<grid(n), 256, 0, stream>>>(ga, ba, ca, q, fo, gx, gy, gz, flags);
        break;
    }
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
