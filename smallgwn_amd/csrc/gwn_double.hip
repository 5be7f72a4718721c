// gwn_double.hip - Real = double objects and batches on gfx950 (gwn_*_object<double, Index>).
//
// Reference: the same templates as the float path instantiated with Real = double
// (include/gwn/gwn_bvh_build.cuh:34-39; detail/gwn_bvh_refit_impl.cuh:23-352;
// detail/gwn_bvh_refit_moment_impl.cuh:17-279; detail/gwn_query_winding_impl.cuh:102-443).
//
// A double BVH is the float build's tree over the positions rounded to float (Morton codes, sort,
// LBVH or H-PLOC, collapse), carrying double boxes and double records computed from the double
// positions by a level-synchronous fold (one launch per breadth-first level, deepest first, as the
// float bounds refit does).  Only the tree shape comes from the rounded positions, and every box
// is the exact union of its double triangles, so queries are exact to double rounding.  Moments
// are the float path's formulas in double (gwn_moment_math.hpp), refit level by level.  The exact
// batch sums the reference's Van Oosterom-Strackee solid angle in double; the Taylor batch is the
// per-lane traversal with the reference's acceptance test and far-field terms in double.
//
// Arithmetic is unfused, like the reference's double path.
#include "gwn_core.hpp"
#include "gwn_moment_math.hpp"

#include <cmath>

#pragma clang fp contract(off)

namespace gwn {
namespace {

using namespace moment_math;

constexpr int k_block = 256;
constexpr double k_pi_d = 3.14159265358979323846;

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; } // std::min(a, b)
__device__ inline double dmax(double a, double b) { return (a < b) ? b : a; } // std::max(a, b)

__global__ __launch_bounds__(k_block) void narrow_positions_kernel(
    double const *__restrict__ xd, double const *__restrict__ yd, double const *__restrict__ zd, float *__restrict__ x,
    float *__restrict__ y, float *__restrict__ z, std::size_t n
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    x[i] = static_cast<float>(xd[i]);
    y[i] = static_cast<float>(yd[i]);
    z[i] = static_cast<float>(zd[i]);
}

// Leaf-ordered double records from the double positions through the primitive order.
__global__ __launch_bounds__(k_block) void records_d_kernel(
    double const *__restrict__ xd, double const *__restrict__ yd, double const *__restrict__ zd, std::size_t nv,
    std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1, std::uint32_t const *__restrict__ i2,
    std::uint32_t const *__restrict__ order, std::size_t nt, tri_record_d *__restrict__ records,
    std::uint32_t *__restrict__ error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= nt)
        return;
    std::uint32_t const p = order[i];
    if (p >= nt) {
        atomicOr(error, 1u);
        return;
    }
    std::uint32_t const a = i0[p], b = i1[p], c = i2[p];
    if (a >= nv || b >= nv || c >= nv) {
        atomicOr(error, 1u);
        return;
    }
    records[i] = tri_record_d{xd[a], yd[a], zd[a], xd[b], yd[b], zd[b], xd[c], yd[c], zd[c]};
}

// Triangle box from the record's vertices (gwn_bvh_triangle_impl.cuh:56-66).
__device__ inline aabb_d triangle_box_d(tri_record_d const &t) {
    return aabb_d{dmin(dmin(t.v0x, t.v1x), t.v2x), dmin(dmin(t.v0y, t.v1y), t.v2y), dmin(dmin(t.v0z, t.v1z), t.v2z),
                  dmax(dmax(t.v0x, t.v1x), t.v2x), dmax(dmax(t.v0y, t.v1y), t.v2y), dmax(dmax(t.v0z, t.v1z), t.v2z)};
}

__device__ inline aabb_d union_d(aabb_d const &a, aabb_d const &b) {
    return aabb_d{dmin(a.min_x, b.min_x), dmin(a.min_y, b.min_y), dmin(a.min_z, b.min_z),
                  dmax(a.max_x, b.max_x), dmax(a.max_y, b.max_y), dmax(a.max_z, b.max_z)};
}

__device__ inline bool leaf_box_d(tri_record_d const *tris, std::size_t nt, std::uint64_t ref, aabb_d &box) {
    std::uint64_t const off = ref_offset(ref);
    std::uint32_t const cnt = ref_count(ref);
    if (cnt == 0 || off + cnt > nt)
        return false;
    box = triangle_box_d(tris[off]);
    for (std::uint32_t k = 1; k < cnt; ++k)
        box = union_d(box, triangle_box_d(tris[off + k]));
    return true;
}

// One breadth-first level of the double fold: each node's double child boxes (leaves from their
// records, internal children from the deeper level's folds) and its own fold in slot order.
__global__ __launch_bounds__(k_block) void bounds_level_d_kernel(
    bvh_node4 const *__restrict__ nodes, bvh_node4_d *__restrict__ nodes_d, std::size_t node_count,
    tri_record_d const *__restrict__ tris, std::size_t nt, std::size_t first, std::size_t count,
    aabb_d *__restrict__ pending, std::uint32_t *error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= count)
        return;
    std::size_t const n = first + i;
    bvh_node4_d out{};
    aabb_d acc{};
    bool init = false;
    for (int s = 0; s < 4; ++s) {
        std::uint64_t const ref = nodes[n].children[s].reference;
        out.children[s].reference = ref;
        if (!ref_valid(ref))
            continue;
        aabb_d box;
        if (ref_internal(ref)) {
            std::uint64_t const c = ref_offset(ref);
            if (c >= node_count || c <= n) {
                atomicOr(error, 1u);
                return;
            }
            box = pending[c];
        } else if (!leaf_box_d(tris, nt, ref, box)) {
            atomicOr(error, 1u);
            return;
        }
        out.children[s].bounds = box;
        acc = init ? union_d(acc, box) : box;
        init = true;
    }
    if (!init) {
        atomicOr(error, 1u);
        return;
    }
    nodes_d[n] = out;
    pending[n] = acc;
}

__global__ void leaf_root_box_d_kernel(tri_record_d const *tris, std::size_t nt, std::uint64_t ref, aabb_d *box,
                                       std::uint32_t *error) {
    if (!leaf_box_d(tris, nt, ref, *box))
        atomicOr(error, 1u);
}

// One breadth-first level of the double moment refit (the float path's moment_refit_node in
// double): each slot's payload, its recentred fields, and the node's total in slot order.
template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_d_kernel(
    bvh_node4_d const *__restrict__ nodes, std::size_t node_count, tri_record_d const *__restrict__ tris,
    std::size_t nt, std::size_t first, std::size_t count, payload_t<double, Order> *__restrict__ pending,
    std::uint32_t *__restrict__ crefs, double *__restrict__ fields, std::uint32_t *error
) {
    constexpr int F = moment_field_count(Order);
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= count)
        return;
    std::size_t const n = first + i;
    double *nf = fields + n * F * 4;
    payload_t<double, Order> total;
    payload_zero(total);
    bool init = false;
    for (int s = 0; s < 4; ++s) {
        bvh_child_d const ch = nodes[n].children[s];
        std::uint64_t const ref = ch.reference;
        if (!ref_valid(ref)) {
            for (int f = 0; f < F; ++f)
                nf[f * 4 + s] = 0.0;
            crefs[n * 4 + s] = k_cref_invalid;
            continue;
        }
        payload_t<double, Order> child;
        if (ref_internal(ref)) {
            std::uint64_t const c = ref_offset(ref);
            if (c >= node_count || c <= n) {
                atomicOr(error, 1u);
                return;
            }
            child = pending[c];
            crefs[n * 4 + s] = static_cast<std::uint32_t>(c);
        } else {
            std::uint64_t const off = ref_offset(ref);
            std::uint32_t const cnt = ref_count(ref);
            if (off + cnt > nt || cnt != 1u) {
                atomicOr(error, 1u);
                return;
            }
            payload_zero(child);
            payload_t<double, Order> tri;
            triangle_payload<Order>(tris[off], tri);
            payload_add(child, tri);
            crefs[n * 4 + s] = k_cref_leaf | static_cast<std::uint32_t>(off);
        }
        write_fields<Order>(nf, s, ch.bounds, child);
        if (init)
            payload_add(total, child);
        else
            total = child;
        init = true;
    }
    pending[n] = total;
}

// The reference's solid angle in double (gwn_query_geometry_impl.cuh:16-45): corners rebuilt from
// (v0, e1 = v1 - v0, e2 = v2 - v0), unit vectors, Van Oosterom-Strackee.
__device__ inline double solid_angle_d(tri_record_d const &t, double qx, double qy, double qz) {
    double const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    double const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    double ax = t.v0x - qx, ay = t.v0y - qy, az = t.v0z - qz;
    double bx = (t.v0x + e1x) - qx, by = (t.v0y + e1y) - qy, bz = (t.v0z + e1z) - qz;
    double cx = (t.v0x + e2x) - qx, cy = (t.v0y + e2y) - qy, cz = (t.v0z + e2z) - qz;
    double const al = sqrt(ax * ax + ay * ay + az * az);
    double const bl = sqrt(bx * bx + by * by + bz * bz);
    double const cl = sqrt(cx * cx + cy * cy + cz * cz);
    if (al == 0.0 || bl == 0.0 || cl == 0.0)
        return 0.0;
    ax /= al; ay /= al; az /= al;
    bx /= bl; by /= bl; bz /= bl;
    cx /= cl; cy /= cl; cz /= cl;
    double const ux = bx - ax, uy = by - ay, uz = bz - az;
    double const vx = cx - ax, vy = cy - ay, vz = cz - az;
    double const num = ax * (uy * vz - uz * vy) + ay * (uz * vx - ux * vz) + az * (ux * vy - uy * vx);
    if (num == 0.0)
        return 0.0;
    double const den = 1.0 + (ax * bx + ay * by + az * bz) + (ax * cx + ay * cy + az * cz) + (bx * cx + by * cy + bz * cz);
    return 2.0 * atan2(num, den);
}

__global__ __launch_bounds__(k_block) void exact_d_kernel(
    tri_record_d const *__restrict__ tris, std::size_t nt, double const *__restrict__ qx,
    double const *__restrict__ qy, double const *__restrict__ qz, double *__restrict__ out, std::size_t nq
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= nq)
        return;
    double const x = qx[i], y = qy[i], z = qz[i];
    double sum = 0.0;
    for (std::size_t k = 0; k < nt; ++k)
        sum += solid_angle_d(tris[k], x, y, z);
    out[i] = sum / (4.0 * k_pi_d);
}

// One accepted far-field expansion in double (gwn_query_winding_impl.cuh:186-325), the float
// path's taylor_term with 1 / sqrt in place of the hardware reciprocal square root.
template <int Order>
__device__ inline double taylor_term_d(double const *f, int s, double qrx, double qry, double qrz, double ql2) {
#define GWN_F(k) f[(k) * 4 + s]
    double const m1 = 1.0 / sqrt(ql2);
    double const m2 = m1 * m1;
    double const m3 = m2 * m1;
    double const qnx = qrx * m1, qny = qry * m1, qnz = qrz * m1;
    double const Nx = GWN_F(4), Ny = GWN_F(5), Nz = GWN_F(6);
    double w = -m2 * (qnx * Nx + qny * Ny + qnz * Nz);
    if constexpr (Order >= 1) {
        double const Txx = GWN_F(7), Tyy = GWN_F(8), Tzz = GWN_F(9);
        double const Sxy = GWN_F(10), Syz = GWN_F(11), Szx = GWN_F(12);
        double const S = qnx * qnx * Txx + qny * qny * Tyy + qnz * qnz * Tzz + qnx * qny * Sxy + qnx * qnz * Szx +
                         qny * qnz * Syz;
        w += m3 * ((Txx + Tyy + Tzz) - 3.0 * S);
    }
    if constexpr (Order >= 2) {
        double const m4 = m2 * m2;
        double const qnx2 = qnx * qnx, qny2 = qny * qny, qnz2 = qnz * qnz;
        double const xxx = GWN_F(13), yyy = GWN_F(14), zzz = GWN_F(15), spn = GWN_F(16);
        double const Axy = GWN_F(17), Axz = GWN_F(18), Ayz = GWN_F(19), Ayx = GWN_F(20), Azx = GWN_F(21),
                     Azy = GWN_F(22);
        double const t0x = Ayx + Azx, t0y = Azy + Axy, t0z = Axz + Ayz;
        double const t1x = qny * Axy + qnz * Axz;
        double const t1y = qnz * Ayz + qnx * Ayx;
        double const t1z = qnx * Azx + qny * Azy;
        w += m4 * (1.5 * (qnx * (3.0 * xxx + t0x) + qny * (3.0 * yyy + t0y) + qnz * (3.0 * zzz + t0z)) -
                   7.5 * (qnx2 * qnx * xxx + qny2 * qny * yyy + qnz2 * qnz * zzz + qnx * qny * qnz * spn +
                          qnx2 * t1x + qny2 * t1y + qnz2 * t1z));
    }
#undef GWN_F
    return w;
}

// Per-lane Taylor traversal in double: children in slot order, the last descended internal child
// entered directly and the others pushed (the float traversal's order); returns true on overflow.
template <int Order, int Cap>
__device__ inline bool taylor_traverse_d(
    std::uint32_t const *__restrict__ crefs, double const *__restrict__ fields, tri_record_d const *__restrict__ tris,
    std::uint32_t root, double x, double y, double z, double beta2, double &omega
) {
    constexpr int F = moment_field_count(Order);
    std::uint32_t stack[Cap];
    int sp = 0;
    bool overflow = false;
    std::uint32_t node = root;
    while (node != k_cref_invalid && !overflow) {
        double const *f = fields + static_cast<std::size_t>(node) * F * 4;
        std::uint32_t next = k_cref_invalid;
        for (int s = 0; s < 4; ++s) {
            std::uint32_t const c = crefs[static_cast<std::size_t>(node) * 4 + s];
            if (c == k_cref_invalid || overflow)
                continue;
            double const qrx = x - f[1 * 4 + s], qry = y - f[2 * 4 + s], qrz = z - f[3 * 4 + s];
            double const ql2 = qrx * qrx + qry * qry + qrz * qrz;
            bool const descend = !(ql2 > 0.0) || ql2 <= f[s] * beta2;
            if (!descend) {
                double const w = taylor_term_d<Order>(f, s, qrx, qry, qrz, ql2);
                if (isfinite(w)) {
                    omega += w;
                    continue;
                }
            }
            if (c & k_cref_leaf) {
                omega += solid_angle_d(tris[c & ~k_cref_leaf], x, y, z);
                continue;
            }
            if (next != k_cref_invalid) {
                if (sp < Cap)
                    stack[sp++] = next;
                else
                    overflow = true;
            }
            next = c;
        }
        if (next == k_cref_invalid && sp > 0)
            next = stack[--sp];
        node = next;
    }
    return overflow;
}

template <int Order, int Cap>
__global__ __launch_bounds__(k_block) void taylor_d_kernel(
    std::uint32_t const *__restrict__ crefs, double const *__restrict__ fields, tri_record_d const *__restrict__ tris,
    std::uint32_t root, double const *__restrict__ qx, double const *__restrict__ qy, double const *__restrict__ qz,
    double *__restrict__ out, std::size_t n, double beta2
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    double omega = 0.0;
    bool const overflow = taylor_traverse_d<Order, Cap>(crefs, fields, tris, root, qx[i], qy[i], qz[i], beta2, omega);
    out[i] = overflow ? __builtin_nan("") : omega / (4.0 * k_pi_d);
}

__global__ __launch_bounds__(k_block) void leaf_root_d_kernel(
    tri_record_d const *__restrict__ tris, std::uint64_t off, std::uint32_t cnt, double const *__restrict__ qx,
    double const *__restrict__ qy, double const *__restrict__ qz, double *__restrict__ out, std::size_t n
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    double sum = 0.0;
    for (std::uint32_t k = 0; k < cnt; ++k)
        sum += solid_angle_d(tris[off + k], qx[i], qy[i], qz[i]);
    out[i] = sum / (4.0 * k_pi_d);
}

} // namespace

void narrow_positions(gwn_geometry_handle &g, hipStream_t stream) {
    if (g.vertex_count == 0)
        return;
    narrow_positions_kernel<<<grid_for(g.vertex_count, k_block), k_block, 0, stream>>>(
        g.xd.ptr, g.yd.ptr, g.zd.ptr, g.x.ptr, g.y.ptr, g.z.ptr, g.vertex_count);
    GWN_LAUNCH_CHECK();
}

void refit_bounds_double(gwn_geometry_handle const &g, gwn_bvh_handle &bvh, hipStream_t stream) {
    std::size_t const T = bvh.triangle_count, N = bvh.node_count;
    if (T == 0)
        return;
    if (bvh.triangles_d.count != T)
        bvh.triangles_d.allocate(T, stream);
    if (bvh.nodes_d.count != N)
        bvh.nodes_d.allocate(N, stream);
    device_buffer<std::uint32_t> error(1, stream);
    fill_u32(error.ptr, 0u, 1, stream);
    records_d_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
        g.xd.ptr, g.yd.ptr, g.zd.ptr, g.vertex_count, g.i0.ptr, g.i1.ptr, g.i2.ptr, bvh.primitive_indices.ptr, T,
        bvh.triangles_d.ptr, error.ptr);
    GWN_LAUNCH_CHECK();
    device_buffer<aabb_d> boxes(N > 0 ? N : 1, stream);
    if (ref_internal(bvh.root.reference)) {
        auto const &ls = bvh.level_start;
        if (ls.size() < 2 || ls.front() != 0u || ls.back() != N)
            throw_internal("bvh.build.double: BVH level table does not cover its nodes.");
        for (std::size_t d = ls.size() - 1; d-- > 0;) {
            std::size_t const first = ls[d], count = ls[d + 1] - ls[d];
            if (count == 0)
                continue;
            bounds_level_d_kernel<<<grid_for(count, k_block), k_block, 0, stream>>>(
                bvh.nodes.ptr, bvh.nodes_d.ptr, N, bvh.triangles_d.ptr, T, first, count, boxes.ptr, error.ptr);
            GWN_LAUNCH_CHECK();
        }
    } else {
        leaf_root_box_d_kernel<<<1, 1, 0, stream>>>(bvh.triangles_d.ptr, T, bvh.root.reference, boxes.ptr, error.ptr);
        GWN_LAUNCH_CHECK();
    }
    std::uint32_t host_error = 0;
    aabb_d root_box{};
    GWN_HIP(hipMemcpyAsync(&host_error, error.ptr, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipMemcpyAsync(&root_box, boxes.ptr, sizeof(aabb_d), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.build.double: Double-precision bounds found an invalid tree.");
    bvh.root_d.bounds = root_box;
    bvh.root_d.reference = bvh.root.reference;
}

void refit_moment_double(gwn_bvh_handle const &bvh, gwn_moment_handle &m, hipStream_t stream) {
    std::size_t const N = ref_internal(bvh.root.reference) ? bvh.node_count : 0;
    m.node_count = bvh.node_count;
    m.width = bvh.width;
    m.real_bytes = 8;
    m.bvh_revision = bvh.revision;
    if (N == 0)
        return;
    auto const &ls = bvh.level_start;
    if (ls.size() < 2 || ls.front() != 0u || ls.back() != N)
        throw_internal("bvh.refit.moment.double: BVH level table does not cover its nodes.");
    int const F = moment_field_count(m.order);
    m.crefs_d.allocate(N * 4, stream);
    m.fields_d.allocate(N * static_cast<std::size_t>(F) * 4, stream);
    std::size_t const pbytes = m.order == 0 ? sizeof(payload_t<double, 0>)
                                            : (m.order == 1 ? sizeof(payload_t<double, 1>) : sizeof(payload_t<double, 2>));
    device_buffer<std::uint8_t> pending(N * pbytes, stream);
    device_buffer<std::uint32_t> error(1, stream);
    fill_u32(error.ptr, 0u, 1, stream);
    for (std::size_t d = ls.size() - 1; d-- > 0;) {
        std::size_t const first = ls[d], count = ls[d + 1] - ls[d];
        if (count == 0)
            continue;
        unsigned const grid = grid_for(count, k_block);
#define GWN_MOMENT_D(O)                                                                                       \
    moment_level_d_kernel<O><<<grid, k_block, 0, stream>>>(                                                  \
        bvh.nodes_d.ptr, N, bvh.triangles_d.ptr, bvh.triangle_count, first, count,                          \
        reinterpret_cast<payload_t<double, O> *>(pending.ptr), m.crefs_d.ptr, m.fields_d.ptr, error.ptr)
        switch (m.order) {
        case 0: GWN_MOMENT_D(0); break;
        case 1: GWN_MOMENT_D(1); break;
        default: GWN_MOMENT_D(2); break;
        }
#undef GWN_MOMENT_D
        GWN_LAUNCH_CHECK();
    }
    std::uint32_t host_error = 0;
    GWN_HIP(hipMemcpyAsync(&host_error, error.ptr, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.refit.moment.double: Level moment refit found an invalid tree.");
}

void winding_exact_double(gwn_bvh_handle const &bvh, double const *qx, double const *qy, double const *qz,
                          double *out, std::size_t n, hipStream_t stream) {
    if (n == 0)
        return;
    exact_d_kernel<<<grid_for(n, k_block), k_block, 0, stream>>>(bvh.triangles_d.ptr, bvh.triangle_count, qx, qy, qz,
                                                                   out, n);
    GWN_LAUNCH_CHECK();
}

void winding_taylor_double(gwn_bvh_handle const &bvh, gwn_moment_handle const &m, double const *qx, double const *qy,
                           double const *qz, double *out, std::size_t n, double accuracy_scale, hipStream_t stream) {
    if (n == 0)
        return;
    unsigned const grid = grid_for(n, k_block);
    if (ref_leaf(bvh.root.reference)) {
        leaf_root_d_kernel<<<grid, k_block, 0, stream>>>(bvh.triangles_d.ptr, ref_offset(bvh.root.reference),
                                                         ref_count(bvh.root.reference), qx, qy, qz, out, n);
        GWN_LAUNCH_CHECK();
        return;
    }
    double const beta2 = accuracy_scale * accuracy_scale;
    std::uint32_t const root = static_cast<std::uint32_t>(ref_offset(bvh.root.reference));
    int const need = static_cast<int>(bvh.internal_stack_bound);
#define GWN_TAYLOR_D(O, C)                                                                                     \
    taylor_d_kernel<O, C><<<grid, k_block, 0, stream>>>(m.crefs_d.ptr, m.fields_d.ptr, bvh.triangles_d.ptr, root, \
                                                       qx, qy, qz, out, n, beta2)
#define GWN_TAYLOR_D_CAP(O)                                                                                    \
    if (need <= 32)                                                                                            \
        GWN_TAYLOR_D(O, 32);                                                                                   \
    else if (need <= 64)                                                                                       \
        GWN_TAYLOR_D(O, 64);                                                                                   \
    else if (need <= 128)                                                                                      \
        GWN_TAYLOR_D(O, 128);                                                                                  \
    else                                                                                                       \
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
    switch (m.order) {
    case 0: GWN_TAYLOR_D_CAP(0) break;
    case 1: GWN_TAYLOR_D_CAP(1) break;
    default: GWN_TAYLOR_D_CAP(2) break;
    }
#undef GWN_TAYLOR_D_CAP
#undef GWN_TAYLOR_D
    GWN_LAUNCH_CHECK();
}

} // namespace gwn
