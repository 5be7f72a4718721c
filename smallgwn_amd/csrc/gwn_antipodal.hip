// gwn_antipodal.hip - Antipodal winding number and gradient batches on gfx950.
//
// Reference: include/gwn/detail/gwn_query_winding_antipodal_impl.cuh:55-552,
// include/gwn/gwn_query.cuh:447-640.
//
// W(q) = (signed crossings of the positive coordinate ray from q with the triangles) + (signed
// spherical areas of the boundary edges seen from q) / 4 pi.  A projected edge or a triangle
// through q makes the axis singular; both terms then restart on the next axis, in the order Z, X,
// Y, and a query singular on all three gets NaN.  The gradient uses the boundary term only.
//
// One query per lane.  Crossings: a stack traversal of the canonical BVH that keeps children
// whose box the axis ray enters; the sum of +-1 is exact in any order.  Boundary: the block
// stages the chain's edges through LDS in 256-edge tiles (endpoints gathered from the vertex
// arrays once per block) and every lane that still needs the current axis walks the tile in
// chain order, with the reference's Kahan sum (plain) or FP64 terms and sums (precision split).
// A second or third pass over the chain runs only when a lane of the block needs another axis.
//
// The classification and crossing arithmetic is unfused, so branch decisions equal the oracle's.
#include "gwn_core.hpp"
#include "../../include/gwn/detail/gwn_device_antipodal.hpp"

#pragma clang fp contract(off)

namespace gwn {
namespace {

using namespace ::gwn::detail::antipodal;
constexpr int k_block = 256;
// One axis pass over the whole chain for the lanes with `need`; returns each lane's accumulation.
template <int Prec, bool Grad>
__device__ inline boundary_acc<Prec, Grad> boundary_pass(
    boundary_view const &bv, staged_edge *tile, v3 const &q, int axis, bool need
) {
    boundary_acc<Prec, Grad> acc;
    for (std::size_t base = 0; base < bv.count; base += k_block) {
        std::size_t const cnt = (bv.count - base) < k_block ? (bv.count - base) : k_block;
        __syncthreads();
        if (threadIdx.x < cnt) {
            std::size_t const i = base + threadIdx.x;
            std::uint32_t const a = bv.start[i], b = bv.end[i];
            tile[threadIdx.x] = staged_edge{bv.vx[a], bv.vy[a], bv.vz[a], bv.vx[b], bv.vy[b], bv.vz[b],
                                            static_cast<float>(bv.mult[i])};
        }
        __syncthreads();
        if (need && !acc.singular) {
            for (std::size_t k = 0; k < cnt; ++k) {
                acc.add(tile[k], q, axis);
                if (acc.singular)
                    break;
            }
        }
    }
    return acc;
}

template <int Prec, int Cap>
__global__ __launch_bounds__(k_block) void antipodal_kernel(
    crossing_scene sc, boundary_view bv, float const *__restrict__ qx, float const *__restrict__ qy,
    float const *__restrict__ qz, float *__restrict__ out, std::size_t n
) {
    __shared__ staged_edge tile[k_block];
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    bool const active = i < n;
    v3 const q{active ? qx[i] : 0.0f, active ? qy[i] : 0.0f, active ? qz[i] : 0.0f};
    bool pending = active;
    float result = __builtin_nanf("");
    int const axes[3] = {k_axis_z, k_axis_x, k_axis_y};
    for (int ai = 0; ai < 3; ++ai) {
        int const axis = axes[ai];
        float crossings = 0.0f;
        bool need = pending;
        if (need && crossing_count<Cap>(sc, q, axis, crossings))
            need = false; // singular on this axis: skip the boundary, try the next axis
        if (!__syncthreads_or(pending))
            break;
        boundary_acc<Prec, false> const acc = boundary_pass<Prec, false>(bv, tile, q, axis, need);
        if (need && !acc.singular) {
            if constexpr (Prec == GWN_PRECISION_PLAIN)
                result = crossings + acc.sum / (4.0f * k_pi_f);
            else
                result = static_cast<float>(double(crossings) + acc.sum_d / (4.0 * k_pi_d));
            pending = false;
        }
    }
    if (active)
        out[i] = result;
}

template <int Prec>
__global__ __launch_bounds__(k_block) void antipodal_gradient_kernel(
    boundary_view bv, float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    float *__restrict__ gx, float *__restrict__ gy, float *__restrict__ gz, std::size_t n
) {
    __shared__ staged_edge tile[k_block];
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    bool const active = i < n;
    v3 const q{active ? qx[i] : 0.0f, active ? qy[i] : 0.0f, active ? qz[i] : 0.0f};
    bool pending = active;
    float const nan = __builtin_nanf("");
    float g[3] = {nan, nan, nan};
    int const axes[3] = {k_axis_z, k_axis_x, k_axis_y};
    for (int ai = 0; ai < 3; ++ai) {
        if (!__syncthreads_or(pending))
            break;
        boundary_acc<Prec, true> const acc = boundary_pass<Prec, true>(bv, tile, q, axes[ai], pending);
        if (pending && !acc.singular) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if constexpr (Prec == GWN_PRECISION_PLAIN)
                    g[k] = acc.gs[k] / (4.0f * k_pi_f);
                else
                    g[k] = static_cast<float>(acc.gd[k] / (4.0 * k_pi_d));
            }
            pending = false;
        }
    }
    if (active) {
        gx[i] = g[0];
        gy[i] = g[1];
        gz[i] = g[2];
    }
}

template <int Prec, int Cap>
void launch_antipodal(
    crossing_scene const &sc, boundary_view const &bv, float const *qx, float const *qy, float const *qz,
    float *out, std::size_t n, hipStream_t stream
) {
    antipodal_kernel<Prec, Cap><<<grid_for(n, k_block), k_block, 0, stream>>>(sc, bv, qx, qy, qz, out, n);
    GWN_LAUNCH_CHECK();
}

template <int Prec>
void launch_antipodal_cap(
    gwn_bvh_handle const &bvh, crossing_scene const &sc, boundary_view const &bv, float const *qx,
    float const *qy, float const *qz, float *out, std::size_t n, hipStream_t stream
) {
    int const need = static_cast<int>(bvh.internal_stack_bound);
    if (need <= 32)
        launch_antipodal<Prec, 32>(sc, bv, qx, qy, qz, out, n, stream);
    else if (need <= 64)
        launch_antipodal<Prec, 64>(sc, bv, qx, qy, qz, out, n, stream);
    else if (need <= 128)
        launch_antipodal<Prec, 128>(sc, bv, qx, qy, qz, out, n, stream);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
}

boundary_view make_view(gwn_geometry_handle const &g, gwn_boundary_handle const &b) {
    return boundary_view{g.x.ptr, g.y.ptr, g.z.ptr, b.start.ptr, b.end.ptr, b.multiplicity.ptr, b.edge_count};
}

} // namespace

void winding_antipodal(
    gwn_geometry_handle const &geometry, gwn_bvh_handle const &bvh, gwn_boundary_handle const &boundary,
    float const *qx, float const *qy, float const *qz, float *out, std::size_t n, int precision, hipStream_t stream
) {
    if (n == 0)
        return;
    crossing_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count, bvh.root.reference};
    boundary_view const bv = make_view(geometry, boundary);
    if (precision == GWN_PRECISION_PLAIN)
        launch_antipodal_cap<GWN_PRECISION_PLAIN>(bvh, sc, bv, qx, qy, qz, out, n, stream);
    else if (precision == GWN_PRECISION_SPLIT)
        launch_antipodal_cap<GWN_PRECISION_SPLIT>(bvh, sc, bv, qx, qy, qz, out, n, stream);
    else
        throw_invalid("Antipodal winding supports the plain and split precision modes.");
}

void winding_gradient_antipodal(
    gwn_geometry_handle const &geometry, gwn_boundary_handle const &boundary, float const *qx, float const *qy,
    float const *qz, float *gx, float *gy, float *gz, std::size_t n, int precision, hipStream_t stream
) {
    if (n == 0)
        return;
    boundary_view const bv = make_view(geometry, boundary);
    if (precision == GWN_PRECISION_PLAIN)
        antipodal_gradient_kernel<GWN_PRECISION_PLAIN><<<grid_for(n, k_block), k_block, 0, stream>>>(
            bv, qx, qy, qz, gx, gy, gz, n);
    else if (precision == GWN_PRECISION_SPLIT)
        antipodal_gradient_kernel<GWN_PRECISION_SPLIT><<<grid_for(n, k_block), k_block, 0, stream>>>(
            bv, qx, qy, qz, gx, gy, gz, n);
    else
        throw_invalid("Antipodal gradient supports the plain and split precision modes.");
    GWN_LAUNCH_CHECK();
}

} // namespace gwn
