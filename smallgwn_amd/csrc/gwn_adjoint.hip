// gwn_adjoint.hip - adjoint of Taylor winding-number losses with respect to vertex positions.
//
// No reference counterpart: the reference has no derivative with respect to the mesh.  For
// L = sum_q u_q w_q (u_q = dL/dw_q from the caller), w the Taylor winding of orders 0 or 1 with
// the forward's accept/descend decisions:
//
//  1. re-traversal (gwn_winding.hip, taylor_adjoint_kernel): every accepted child slot collects
//     u/(4 pi) (dw/dP, dw/dN, dw/dT) and every near triangle u/(4 pi) dW/d(a, b, c);
//  2. slot_adjoint_kernel: a slot's fields are P = AP/A, N and T_ij = (N c)_ij - P_j N_i
//     (symmetrised off the diagonal) of the slot's additive payload (A, AP, N, N c) -- the field
//     adjoints become payload adjoints (16 floats per slot);
//  3. push_kernel: a slot's payload is the sum of the payloads of the triangles below it, so each
//     triangle's payload adjoint is the sum of the slot adjoints on its leaf-to-root path (parent
//     links); the triangle payload A = |N|, N = (b - a) x (c - a) / 2, AP = A c, (N c) with
//     c = (a + b + c) / 3 then gives the vertex adjoints, added with the near-field ones into the
//     caller's vertex arrays (atomics).
//
// Queries are visited in Hilbert order for large batches.  Sums use FP32 atomics, so the result is
// reproducible to rounding, not bit for bit.
#include "gwn_core.hpp"

namespace gwn {
namespace {

constexpr int k_block = 256;
constexpr std::uint32_t k_invalid = 0xffffffffu;

// Field adjoints (12 per slot: dP, dN, dT) -> payload adjoints (16 per slot: dA, dAP, dN, d(N c)).
template <int Order>
__global__ __launch_bounds__(k_block) void slot_adjoint_kernel(
    float4 const *__restrict__ records, float const *__restrict__ slot_area, std::size_t slots,
    float const *__restrict__ dslot, float *__restrict__ dpay
) {
    constexpr int RV = moment_record_vec4(Order);
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= slots)
        return;
    std::size_t const node = i >> 2;
    int const s = static_cast<int>(i & 3);
    float const *fields = reinterpret_cast<float const *>(records + node * RV + 1);
#define GWN_F(k) fields[(k) * 4 + s]
    float const P[3] = {GWN_F(1), GWN_F(2), GWN_F(3)};
    float const N[3] = {GWN_F(4), GWN_F(5), GWN_F(6)};
#undef GWN_F
    float const *g = dslot + i * 12;
    float dP[3] = {g[0], g[1], g[2]};
    float dN[3] = {g[3], g[4], g[5]};
    float dn1[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if constexpr (Order >= 1) {
        float const *dT = g + 6;
        // T0..2 = n1[ii] - P_i N_i
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dn1[4 * k] += dT[k];
            dP[k] -= dT[k] * N[k];
            dN[k] -= dT[k] * P[k];
        }
        // T3 = n1[xy] - Py Nx + n1[yx] - Px Ny;  T4 = (y, z);  T5 = (z, x)
        int const pa[3] = {0, 1, 2}, pb[3] = {1, 2, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int const a = pa[k], b = pb[k];
            float const t = dT[3 + k];
            dn1[3 * a + b] += t;
            dn1[3 * b + a] += t;
            dP[b] -= t * N[a];
            dN[a] -= t * P[b];
            dP[a] -= t * N[b];
            dN[b] -= t * P[a];
        }
    }
    float const A = slot_area[i];
    float dA = 0.0f, dAP[3] = {0.0f, 0.0f, 0.0f};
    if (A > 0.0f) {
        // P = AP / A
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dAP[k] = dP[k] / A;
            dA -= dP[k] * P[k] / A;
        }
    }
    float *o = dpay + i * 16;
    o[0] = dA;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o[1 + k] = dAP[k];
        o[4 + k] = dN[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k)
        o[7 + k] = dn1[k];
}

// leaf_slot[record] = node * 4 + slot of the leaf child that holds the record.
__global__ __launch_bounds__(k_block) void leaf_slot_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, std::size_t nt, std::uint32_t *__restrict__ leaf_slot
) {
    std::size_t const n = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (n >= node_count)
        return;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        std::uint64_t const ref = nodes[n].children[s].reference;
        if (!ref_leaf(ref))
            continue;
        std::uint64_t const off = ref_offset(ref);
        std::uint32_t const cnt = ref_count(ref);
        for (std::uint32_t k = 0; k < cnt; ++k)
            if (off + k < nt)
                leaf_slot[off + k] = static_cast<std::uint32_t>(n * 4 + static_cast<std::size_t>(s));
    }
}

// Per triangle record: payload adjoint summed over the leaf-to-root path, chained to the three
// vertices, plus the near-field vertex adjoints; added to the caller's vertex arrays.
__global__ __launch_bounds__(k_block) void push_kernel(
    tri_record const *__restrict__ tris, std::uint32_t const *__restrict__ prim, std::uint32_t const *__restrict__ i0,
    std::uint32_t const *__restrict__ i1, std::uint32_t const *__restrict__ i2, std::size_t nt,
    std::uint32_t const *__restrict__ leaf_slot, std::uint32_t const *__restrict__ parent,
    float const *__restrict__ dpay, float const *__restrict__ dtri, float *dvx, float *dvy, float *dvz
) {
    std::size_t const k = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (k >= nt)
        return;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        acc[j] = 0.0f;
    if (leaf_slot != nullptr) {
        std::uint32_t link = leaf_slot[k];
        while (link != k_invalid) {
            float const *d = dpay + static_cast<std::size_t>(link) * 16;
#pragma unroll
            for (int j = 0; j < 16; ++j)
                acc[j] += d[j];
            link = parent[link >> 2];
        }
    }
    tri_record const t = tris[k];
    float const a[3] = {t.v0x, t.v0y, t.v0z}, b[3] = {t.v1x, t.v1y, t.v1z}, c[3] = {t.v2x, t.v2y, t.v2z};
    float const e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float const e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float const n[3] = {0.5f * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5f * (e1[2] * e2[0] - e1[0] * e2[2]),
                        0.5f * (e1[0] * e2[1] - e1[1] * e2[0])};
    float const cen[3] = {(a[0] + b[0] + c[0]) / 3.0f, (a[1] + b[1] + c[1]) / 3.0f, (a[2] + b[2] + c[2]) / 3.0f};
    float const area = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    float gN[3], gc[3];
    float darea = acc[0] + acc[1] * cen[0] + acc[2] * cen[1] + acc[3] * cen[2];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        gN[j] = acc[4 + j];
        gc[j] = acc[1 + j] * area;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            gN[r] += acc[7 + 3 * r + col] * cen[col];
            gc[col] += acc[7 + 3 * r + col] * n[r];
        }
    if (area > 0.0f) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            gN[j] += darea * n[j] / area;
    }
    // N = (a x b + b x c + c x a) / 2:  d(gN . N)/da = (b - c) x gN / 2, cyclically.
    float const bc[3] = {b[0] - c[0], b[1] - c[1], b[2] - c[2]};
    float const ca[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float const ab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    float const *near = dtri + k * 9;
    std::uint32_t const p = prim[k];
    std::uint32_t const ids[3] = {i0[p], i1[p], i2[p]};
    float const *edges[3] = {bc, ca, ab};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float const *e = edges[v];
        float const gx = 0.5f * (e[1] * gN[2] - e[2] * gN[1]) + gc[0] / 3.0f + near[3 * v + 0];
        float const gy = 0.5f * (e[2] * gN[0] - e[0] * gN[2]) + gc[1] / 3.0f + near[3 * v + 1];
        float const gz = 0.5f * (e[0] * gN[1] - e[1] * gN[0]) + gc[2] / 3.0f + near[3 * v + 2];
        atomicAdd(dvx + ids[v], gx);
        atomicAdd(dvy + ids[v], gy);
        atomicAdd(dvz + ids[v], gz);
    }
}

} // namespace

void winding_taylor_adjoint(
    gwn_geometry_handle const &g, gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx,
    float const *qy, float const *qz, float const *u, std::size_t n, float accuracy_scale, float *dvx, float *dvy,
    float *dvz, float *dqx, float *dqy, float *dqz, hipStream_t stream
) {
    std::size_t const T = bvh.triangle_count;
    if (n == 0 || T == 0)
        return;
    bool const internal = ref_internal(bvh.root.reference);
    std::size_t const N = internal ? bvh.node_count : 0;
    device_buffer<float> dslot(N * 4 * 12 + 1, stream), dpay(N * 4 * 16 + 1, stream), dtri(T * 9, stream);
    device_buffer<std::uint32_t> leaf_slot(internal ? T : 1, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(dslot.ptr), 0u, N * 4 * 12 + 1, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(dtri.ptr), 0u, T * 9, stream);
    device_buffer<std::uint32_t> order;
    if (internal && n >= (1u << 16)) {
        order.allocate(n, stream);
        query_order(bvh, qx, qy, qz, order.ptr, n, stream);
    }
    taylor_adjoint_traverse(bvh, m, qx, qy, qz, u, order.ptr, n, accuracy_scale, dslot.ptr, dtri.ptr, dqx, dqy,
                            dqz, stream);
    if (internal) {
        std::size_t const slots = N * 4;
        if (m.order == 0)
            slot_adjoint_kernel<0><<<grid_for(slots, k_block), k_block, 0, stream>>>(
                m.records.ptr, m.slot_area.ptr, slots, dslot.ptr, dpay.ptr);
        else
            slot_adjoint_kernel<1><<<grid_for(slots, k_block), k_block, 0, stream>>>(
                m.records.ptr, m.slot_area.ptr, slots, dslot.ptr, dpay.ptr);
        GWN_LAUNCH_CHECK();
        leaf_slot_kernel<<<grid_for(N, k_block), k_block, 0, stream>>>(bvh.nodes.ptr, N, T, leaf_slot.ptr);
        GWN_LAUNCH_CHECK();
    }
    push_kernel<<<grid_for(T, k_block), k_block, 0, stream>>>(
        bvh.triangles.ptr, bvh.primitive_indices.ptr, g.i0.ptr, g.i1.ptr, g.i2.ptr, T,
        internal ? leaf_slot.ptr : nullptr, bvh.parent.ptr, dpay.ptr, dtri.ptr, dvx, dvy, dvz);
    GWN_LAUNCH_CHECK();
}

} // namespace gwn
