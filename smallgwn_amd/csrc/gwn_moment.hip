// gwn_moment.hip - bottom-up Taylor moment refit (orders 0, 1, 2) on gfx950.
//
// Reference: include/gwn/detail/gwn_bvh_refit_moment_impl.cuh:17-279 and
// include/gwn/detail/gwn_bvh_refit_common.cuh:14-609.
//
// One kernel.  A thread starts at every wide node without internal children; it computes the
// additive payload of each child slot (leaf payloads straight from the triangle records, so
// leaves need no pending storage), writes the node's recentred coefficients, stores the node's
// total payload and climbs: the last internal child to arrive at a parent (acquire/release
// arrival counter) continues with the parent.  Children are always summed in slot order from the
// first valid slot, exactly as the reference's combine_node, so the result does not depend on
// arrival order.  Arrival counters reset themselves, so a refit needs no clearing pass.
//
// Output: one traversal record per node -- a uint4 of compact child references followed by the
// reference's field-SoA gwn_bvh_moment_node<4, Order, float> coefficients as float4 (one float
// per child slot), padded to 128 / 256 / 384 bytes.
#include "gwn_core.hpp"

// Bit-for-bit agreement with the CPU oracle requires unfused float arithmetic.
#pragma clang fp contract(off)

namespace gwn {
namespace {

constexpr int k_block = 128;
constexpr std::uint32_t k_invalid = 0xffffffffu;

template <int Order> struct payload {
    static constexpr int count = Order == 0 ? 7 : (Order == 1 ? 16 : 34);
    float v[count];
};

__device__ inline float fmax_ref(float a, float b) { return (a < b) ? b : a; } // std::max(a, b)

template <int Order> __device__ inline void payload_zero(payload<Order> &p) {
#pragma unroll
    for (int i = 0; i < payload<Order>::count; ++i)
        p.v[i] = 0.0f;
}

template <int Order> __device__ inline void payload_add(payload<Order> &d, payload<Order> const &s) {
#pragma unroll
    for (int i = 0; i < payload<Order>::count; ++i)
        d.v[i] += s.v[i];
}

// HDK sorted-vertex split integrals for one axis (gwn_bvh_refit_common.cuh:227-281).
__device__ inline void axis_integrals(
    float const (&verts)[3][3], float const (&P)[3], int va, int vb, int vc, float *out_ii,
    float *out_ij, float *out_ik, int i
) {
    int const j = (i == 2) ? 0 : (i + 1);
    int const k = (j == 2) ? 0 : (j + 1);
    float const oab_i = verts[vb][i] - verts[va][i];
    float const oab_j = verts[vb][j] - verts[va][j];
    float const oab_k = verts[vb][k] - verts[va][k];
    float const oac_i = verts[vc][i] - verts[va][i];
    float const oac_j = verts[vc][j] - verts[va][j];
    float const oac_k = verts[vc][k] - verts[va][k];
    float const ocb_i = verts[vb][i] - verts[vc][i];
    float const ocb_j = verts[vb][j] - verts[vc][j];
    float const ocb_k = verts[vb][k] - verts[vc][k];
    float const t = (oac_i > 0.0f) ? (oab_i / oac_i) : 0.0f;
    float const jdiff = t * oac_j - oab_j;
    float const kdiff = t * oac_k - oab_k;
    float const ca0 = jdiff * oab_k - kdiff * oab_j;
    float const ca1 = kdiff * oab_i;
    float const ca2 = jdiff * oab_i;
    float const cc0 = jdiff * ocb_k - kdiff * ocb_j;
    float const cc1 = kdiff * ocb_i;
    float const cc2 = jdiff * ocb_i;
    float const sa = sqrtf(fmax_ref(ca0 * ca0 + ca1 * ca1 + ca2 * ca2, 0.0f));
    float const sc = sqrtf(fmax_ref(cc0 * cc0 + cc1 * cc1 + cc2 * cc2, 0.0f));
    float const Pai = verts[va][i] - P[i];
    float const Pci = verts[vc][i] - P[i];
    *out_ii = sa * (0.5f * Pai * Pai + float(2.0 / 3.0) * Pai * oab_i + 0.25f * oab_i * oab_i) +
              sc * (0.5f * Pci * Pci + float(2.0 / 3.0) * Pci * ocb_i + 0.25f * ocb_i * ocb_i);
    for (int pass = 0; pass < 2; ++pass) {
        float *out = pass == 0 ? out_ij : out_ik;
        int const jk = pass == 0 ? j : k;
        float const diff = pass == 0 ? jdiff : kdiff;
        if (out == nullptr)
            continue;
        float const obmidj = verts[vb][jk] + 0.5f * diff;
        float const oabmidj = obmidj - verts[va][jk];
        float const ocbmidj = obmidj - verts[vc][jk];
        float const Paj = verts[va][jk] - P[jk];
        float const Pcj = verts[vc][jk] - P[jk];
        *out = sa * (0.5f * Pai * Paj + float(1.0 / 3.0) * Pai * oabmidj +
                     float(1.0 / 3.0) * Paj * oab_i + 0.25f * oab_i * oabmidj) +
               sc * (0.5f * Pci * Pcj + float(1.0 / 3.0) * Pci * ocbmidj +
                     float(1.0 / 3.0) * Pcj * ocb_i + 0.25f * ocb_i * ocbmidj);
    }
}

// Additive moments of one leaf record (gwn_bvh_refit_common.cuh:122-387), on the reference's
// reconstructed vertices b = v0 + (v1 - v0), c = v0 + (v2 - v0).
template <int Order> __device__ inline void triangle_payload(tri_record const &r, payload<Order> &m) {
    float a[3], b[3], c[3];
    {
        float const v0[3] = {r.v0x, r.v0y, r.v0z};
        float const v1[3] = {r.v1x, r.v1y, r.v1z};
        float const v2[3] = {r.v2x, r.v2y, r.v2z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float const e1 = v1[k] - v0[k];
            float const e2 = v2[k] - v0[k];
            a[k] = v0[k];
            b[k] = v0[k] + e1;
            c[k] = v0[k] + e2;
        }
    }
    payload_zero(m);
    float const avx = (a[0] + b[0] + c[0]) / 3.0f;
    float const avy = (a[1] + b[1] + c[1]) / 3.0f;
    float const avz = (a[2] + b[2] + c[2]) / 3.0f;
    float const abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
    float const acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
    float const nx = 0.5f * (aby * acz - abz * acy);
    float const ny = 0.5f * (abz * acx - abx * acz);
    float const nz = 0.5f * (abx * acy - aby * acx);
    float const area2 = nx * nx + ny * ny + nz * nz;
    float const area = sqrtf(fmax_ref(area2, 0.0f));
    m.v[0] = area;
    m.v[1] = avx * area;
    m.v[2] = avy * area;
    m.v[3] = avz * area;
    m.v[4] = nx;
    m.v[5] = ny;
    m.v[6] = nz;
    if constexpr (Order >= 1) {
        m.v[7] = nx * avx;
        m.v[8] = nx * avy;
        m.v[9] = nx * avz;
        m.v[10] = ny * avx;
        m.v[11] = ny * avy;
        m.v[12] = ny * avz;
        m.v[13] = nz * avx;
        m.v[14] = nz * avy;
        m.v[15] = nz * avz;
    }
    if constexpr (Order >= 2) {
        if (area == 0.0f)
            return;
        float const inv_area = 1.0f / area;
        float const nh[3] = {nx * inv_area, ny * inv_area, nz * inv_area};
        float const verts[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
        float const P[3] = {avx, avy, avz};
        int ord[3][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}};
        float span[3];
        for (int axis = 0; axis < 3; ++axis) {
            int t;
            if (verts[0][axis] > verts[1][axis]) { t = ord[axis][0]; ord[axis][0] = ord[axis][1]; ord[axis][1] = t; }
            if (verts[ord[axis][0]][axis] > verts[2][axis]) { t = ord[axis][0]; ord[axis][0] = ord[axis][2]; ord[axis][2] = t; }
            if (verts[ord[axis][1]][axis] > verts[ord[axis][2]][axis]) { t = ord[axis][1]; ord[axis][1] = ord[axis][2]; ord[axis][2] = t; }
            span[axis] = verts[ord[axis][2]][axis] - verts[ord[axis][0]][axis];
        }
        float ixx = 0.0f, iyy = 0.0f, izz = 0.0f, ixy = 0.0f, iyz = 0.0f, izx = 0.0f;
        float const dx = span[0], dy = span[1], dz = span[2];
        if (dx > 0.0f)
            axis_integrals(verts, P, ord[0][0], ord[0][1], ord[0][2], &ixx,
                           (dx >= dy && dy > 0.0f) ? &ixy : nullptr,
                           (dx >= dz && dz > 0.0f) ? &izx : nullptr, 0);
        if (dy > 0.0f)
            axis_integrals(verts, P, ord[1][0], ord[1][1], ord[1][2], &iyy,
                           (dy >= dz && dz > 0.0f) ? &iyz : nullptr,
                           (dx < dy && dx > 0.0f) ? &ixy : nullptr, 1);
        if (dz > 0.0f)
            axis_integrals(verts, P, ord[2][0], ord[2][1], ord[2][2], &izz,
                           (dx < dz && dx > 0.0f) ? &izx : nullptr,
                           (dy < dz && dy > 0.0f) ? &iyz : nullptr, 2);
        float const N[3] = {nx, ny, nz};
        float const I[6] = {ixx, iyy, izz, ixy, iyz, izx};
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            float *rr = m.v + 16 + 6 * comp;
            rr[0] = nh[comp] * I[0] + N[comp] * avx * avx;
            rr[1] = nh[comp] * I[1] + N[comp] * avy * avy;
            rr[2] = nh[comp] * I[2] + N[comp] * avz * avz;
            rr[3] = nh[comp] * I[3] + N[comp] * avx * avy;
            rr[4] = nh[comp] * I[4] + N[comp] * avy * avz;
            rr[5] = nh[comp] * I[5] + N[comp] * avz * avx;
        }
    }
}

// Recentre one child slot into the query fields (gwn_bvh_refit_common.cuh:513-607).
template <int Order>
__device__ inline void write_fields(float *fields, int slot, aabb const &bnd, payload<Order> const &p) {
#define GWN_FIELD(f) fields[(f) * 4 + slot]
    float ax, ay, az;
    float const area = p.v[0];
    if (area > 0.0f) {
        ax = p.v[1] / area;
        ay = p.v[2] / area;
        az = p.v[3] / area;
    } else {
        ax = (bnd.min_x + bnd.max_x) * 0.5f;
        ay = (bnd.min_y + bnd.max_y) * 0.5f;
        az = (bnd.min_z + bnd.max_z) * 0.5f;
    }
    float const dx = fmax_ref(ax - bnd.min_x, bnd.max_x - ax);
    float const dy = fmax_ref(ay - bnd.min_y, bnd.max_y - ay);
    float const dz = fmax_ref(az - bnd.min_z, bnd.max_z - az);
    float const Nx = p.v[4], Ny = p.v[5], Nz = p.v[6];
    GWN_FIELD(0) = dx * dx + dy * dy + dz * dz;
    GWN_FIELD(1) = ax;
    GWN_FIELD(2) = ay;
    GWN_FIELD(3) = az;
    GWN_FIELD(4) = Nx;
    GWN_FIELD(5) = Ny;
    GWN_FIELD(6) = Nz;
    if constexpr (Order >= 1) {
        float const *f = p.v + 7;
        float const nxx = f[0] - ax * Nx;
        float const nyy = f[4] - ay * Ny;
        float const nzz = f[8] - az * Nz;
        float const nxy = f[1] - ay * Nx;
        float const nyx = f[3] - ax * Ny;
        float const nyz = f[5] - az * Ny;
        float const nzy = f[7] - ay * Nz;
        float const nzx = f[6] - ax * Nz;
        float const nxz = f[2] - az * Nx;
        GWN_FIELD(7) = nxx;
        GWN_FIELD(8) = nyy;
        GWN_FIELD(9) = nzz;
        GWN_FIELD(10) = nxy + nyx;
        GWN_FIELD(11) = nyz + nzy;
        GWN_FIELD(12) = nzx + nxz;
    }
    if constexpr (Order >= 2) {
        float const Px = ax, Py = ay, Pz = az;
        float const Nc[3] = {Nx, Ny, Nz};
        float c[3][6];
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            float const *raw = p.v + 16 + 6 * comp;
            float const *F = p.v + 7 + 3 * comp;
            float const Nn = Nc[comp];
            c[comp][0] = raw[0] - 2.0f * Px * F[0] + Px * Px * Nn;
            c[comp][1] = raw[1] - 2.0f * Py * F[1] + Py * Py * Nn;
            c[comp][2] = raw[2] - 2.0f * Pz * F[2] + Pz * Pz * Nn;
            c[comp][3] = raw[3] - Px * F[1] - Py * F[0] + Px * Py * Nn;
            c[comp][4] = raw[4] - Py * F[2] - Pz * F[1] + Py * Pz * Nn;
            c[comp][5] = raw[5] - Pz * F[0] - Px * F[2] + Pz * Px * Nn;
        }
        GWN_FIELD(13) = c[0][0];
        GWN_FIELD(14) = c[1][1];
        GWN_FIELD(15) = c[2][2];
        GWN_FIELD(16) = 2.0f * (c[0][4] + c[1][5] + c[2][3]);
        GWN_FIELD(17) = 2.0f * c[0][3] + c[1][0];
        GWN_FIELD(18) = 2.0f * c[0][5] + c[2][0];
        GWN_FIELD(19) = 2.0f * c[1][4] + c[2][1];
        GWN_FIELD(20) = 2.0f * c[1][3] + c[0][1];
        GWN_FIELD(21) = 2.0f * c[2][5] + c[0][2];
        GWN_FIELD(22) = 2.0f * c[2][4] + c[1][2];
    }
#undef GWN_FIELD
}

// One thread's bottom-up pass: from node `start` when it has no internal children, then up the
// parent links while this thread is the last to arrive at the parent.
template <int Order>
__device__ inline void moment_refit_from(
    std::size_t start, bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, std::uint32_t const *__restrict__ parent,
    std::uint8_t const *__restrict__ internal_children, std::uint32_t *arrivals,
    payload<Order> *pending, float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error
) {
    constexpr int RV = moment_record_vec4(Order);
    if (internal_children[start] != 0)
        return;
    std::size_t n = start;
    for (;;) {
        float4 *rec = records + n * RV;
        float *fields = reinterpret_cast<float *>(rec + 1);
        std::uint32_t cref[4] = {k_invalid, k_invalid, k_invalid, k_invalid};
        payload<Order> total;
        payload_zero(total);
        bool init = false;
        for (int s = 0; s < 4; ++s) {
            bvh_child const ch = nodes[n].children[s];
            std::uint64_t const ref = ch.reference;
            if (!ref_valid(ref))
                continue;
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
        rec[0] = make_float4(__uint_as_float(cref[0]), __uint_as_float(cref[1]),
                             __uint_as_float(cref[2]), __uint_as_float(cref[3]));
        std::uint32_t const link = parent[n];
        if (link == k_invalid)
            return;
        pending[n] = total;
        std::size_t const pn = link >> 2;
        release_agent();
        std::uint32_t const prior =
            __hip_atomic_fetch_add(&arrivals[pn], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prior + 1u < internal_children[pn])
            return;
        acquire_agent();
        arrivals[pn] = 0u;
        n = pn;
    }
}

template <int Order>
__global__ __launch_bounds__(k_block) void moment_refit_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, std::uint32_t const *__restrict__ parent,
    std::uint8_t const *__restrict__ internal_children, std::uint32_t *arrivals,
    payload<Order> *pending, float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error
) {
    std::size_t const start = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (start >= node_count)
        return;
    moment_refit_from<Order>(start, nodes, node_count, tris, tri_count, parent, internal_children, arrivals,
                             pending, records, slot_area, error);
}

// Batched refits: one thread per node of every mesh in the table (mesh found by binary search over
// the node prefixes), so many meshes' refits take two launches instead of four each.
__device__ inline std::uint32_t refit_mesh_of(moment_refit_mesh const *meshes, std::uint32_t count, std::uint64_t g) {
    std::uint32_t lo = 0, hi = count;  // meshes[lo].first_node <= g, and g < meshes[hi].first_node
    while (hi - lo > 1) {
        std::uint32_t const mid = lo + (hi - lo) / 2;
        if (meshes[mid].first_node <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Zeroes what moment_refit_enqueue's fills zero: each node's record and slot areas, its arrival
// counter, and the mesh's error word.
template <int Order>
__global__ __launch_bounds__(k_block) void moment_refit_batch_clear_kernel(
    moment_refit_mesh const *__restrict__ meshes, std::uint32_t count, std::uint64_t total
) {
    std::uint64_t const g = static_cast<std::uint64_t>(blockIdx.x) * k_block + threadIdx.x;
    if (g >= total)
        return;
    moment_refit_mesh const m = meshes[refit_mesh_of(meshes, count, g)];
    std::uint64_t const n = g - m.first_node;
    constexpr int RV = moment_record_vec4(Order);
    for (int k = 0; k < RV; ++k)
        m.records[n * RV + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int s = 0; s < 4; ++s)
        m.slot_area[n * 4 + s] = 0.0f;
    m.arrivals[n] = 0u;
    if (n == 0)
        m.arrivals[m.node_count] = 0u;
}

template <int Order>
__global__ __launch_bounds__(k_block) void moment_refit_batch_kernel(
    moment_refit_mesh const *__restrict__ meshes, std::uint32_t count, std::uint64_t total
) {
    std::uint64_t const g = static_cast<std::uint64_t>(blockIdx.x) * k_block + threadIdx.x;
    if (g >= total)
        return;
    moment_refit_mesh const m = meshes[refit_mesh_of(meshes, count, g)];
    moment_refit_from<Order>(g - m.first_node, m.nodes, m.node_count, m.tris, m.tri_count, m.parent,
                             m.internal_children, m.arrivals, static_cast<payload<Order> *>(m.pending), m.records,
                             m.slot_area, m.arrivals + m.node_count);
}

template <int Order>
void launch_refit_batch(moment_refit_mesh const *meshes, std::uint32_t count, std::uint64_t total, hipStream_t stream) {
    unsigned const grid = grid_for(total, k_block);
    moment_refit_batch_clear_kernel<Order><<<grid, k_block, 0, stream>>>(meshes, count, total);
    GWN_LAUNCH_CHECK();
    moment_refit_batch_kernel<Order><<<grid, k_block, 0, stream>>>(meshes, count, total);
    GWN_LAUNCH_CHECK();
}

template <int Order>
void launch_refit(gwn_bvh_handle const &bvh, gwn_moment_handle &m, hipStream_t stream) {
    std::size_t const N = bvh.node_count;
    constexpr int RV = moment_record_vec4(Order);
    m.records.allocate(N * RV, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(m.records.ptr), 0u, N * RV * 4, stream);
    m.slot_area.allocate(N * 4, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(m.slot_area.ptr), 0u, N * 4, stream);
    device_buffer<payload<Order>> pending(N, stream);
    // Per-call arrival counters (plus the error word) so refits of several orders may run
    // concurrently on one BVH.
    device_buffer<std::uint32_t> arrivals(N + 1, stream);
    fill_u32(arrivals.ptr, 0u, N + 1, stream);
    std::uint32_t *error = arrivals.ptr + N;
    moment_refit_kernel<Order><<<grid_for(N, k_block), k_block, 0, stream>>>(
        bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, bvh.parent.ptr,
        bvh.internal_children.ptr, arrivals.ptr, pending.ptr, m.records.ptr, m.slot_area.ptr, error
    );
    GWN_LAUNCH_CHECK();
    std::uint32_t host_error = 0;
    GWN_HIP(hipMemcpyAsync(&host_error, error, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.refit.moment: Bottom-up moment refit did not converge.");
}

template <int Order>
void enqueue_refit(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream) {
    std::size_t const N = bvh.node_count;
    constexpr int RV = moment_record_vec4(Order);
    fill_u32(reinterpret_cast<std::uint32_t *>(m.records.ptr), 0u, N * RV * 4, stream);
    fill_u32(reinterpret_cast<std::uint32_t *>(m.slot_area.ptr), 0u, N * 4, stream);
    fill_u32(w.arrivals.ptr, 0u, N + 1, stream);
    moment_refit_kernel<Order><<<grid_for(N, k_block), k_block, 0, stream>>>(
        bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, bvh.parent.ptr, bvh.internal_children.ptr,
        w.arrivals.ptr, reinterpret_cast<payload<Order> *>(w.pending.ptr), m.records.ptr, m.slot_area.ptr,
        w.arrivals.ptr + N);
    GWN_LAUNCH_CHECK();
}

std::size_t payload_bytes(int order) {
    return order == 0 ? sizeof(payload<0>) : (order == 1 ? sizeof(payload<1>) : sizeof(payload<2>));
}

} // namespace

void refit_moment_prepare(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream) {
    std::size_t const N = ref_internal(bvh.root.reference) ? bvh.node_count : 0;
    m.node_count = bvh.node_count;
    if (N == 0)
        return;
    m.records.allocate(N * static_cast<std::size_t>(moment_record_vec4(m.order)), stream);
    m.slot_area.allocate(N * 4, stream);
    w.pending.allocate(N * payload_bytes(m.order), stream);
    w.arrivals.allocate(N + 1, stream);
}

void refit_moment_enqueue(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream) {
    m.node_count = bvh.node_count;
    m.bvh_revision = bvh.revision;
    if (!ref_internal(bvh.root.reference) || bvh.node_count == 0)
        return;
    switch (m.order) {
    case 0: enqueue_refit<0>(bvh, m, w, stream); break;
    case 1: enqueue_refit<1>(bvh, m, w, stream); break;
    default: enqueue_refit<2>(bvh, m, w, stream); break;
    }
}

void refit_moment_batch(moment_refit_mesh const *meshes, std::uint32_t count, std::uint64_t total_nodes, int order,
                        hipStream_t stream) {
    if (count == 0 || total_nodes == 0)
        return;
    switch (order) {
    case 0: launch_refit_batch<0>(meshes, count, total_nodes, stream); break;
    case 1: launch_refit_batch<1>(meshes, count, total_nodes, stream); break;
    case 2: launch_refit_batch<2>(meshes, count, total_nodes, stream); break;
    default: throw_invalid("Taylor order must be 0, 1, or 2.");
    }
}

void refit_moment(gwn_bvh_handle const &bvh, gwn_moment_handle &m, hipStream_t stream) {
    m.node_count = bvh.node_count;
    m.bvh_revision = bvh.revision;
    if (!ref_internal(bvh.root.reference) || bvh.node_count == 0)
        return; // a leaf root has no moment nodes (gwn_bvh.cuh:290-291)
    switch (m.order) {
    case 0: launch_refit<0>(bvh, m, stream); break;
    case 1: launch_refit<1>(bvh, m, stream); break;
    case 2: launch_refit<2>(bvh, m, stream); break;
    default: throw_invalid("Taylor order must be 0, 1, or 2.");
    }
}

} // namespace gwn
