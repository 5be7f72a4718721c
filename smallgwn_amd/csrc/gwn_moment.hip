// gwn_moment.hip - bottom-up Taylor moment refit (orders 0, 1, 2) on gfx950.
//
// Reference: include/gwn/detail/gwn_bvh_refit_moment_impl.cuh:17-279 and
// include/gwn/detail/gwn_bvh_refit_common.cuh:14-609.
//
// One launch per breadth-first level of the wide tree, deepest first.  A thread takes one node
// of the level; it computes the additive payload of each child slot (leaf payloads straight from
// the triangle records; an internal child's total was stored by the previous, deeper launch),
// writes the node's recentred coefficients and stores the node's total.  Children are always
// summed in slot order from the first valid slot, exactly as the reference's combine_node, and
// the kernel boundary between levels is the only ordering needed: no arrival counters, atomics or
// cross-workgroup fences.
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

// Second-moment integrals of one triangle along axis i (HDK's sorted-vertex split,
// gwn_bvh_refit_common.cuh:227-281).  The vertices are ordered along axis i as lo, mid, hi; the
// line through `mid` perpendicular to axis i cuts the triangle into a lower piece (lo side) and an
// upper piece (hi side), each integrated in closed form about the triangle centroid P.  Writes the
// (i, i) integral and, when requested, the (i, j) and (i, k) cross integrals.
__device__ inline void axis_integrals(
    float const (&verts)[3][3], float const (&P)[3], int lo, int mid, int hi, float *out_ii,
    float *out_ij, float *out_ik, int i
) {
    int const j = (i == 2) ? 0 : (i + 1);
    int const k = (j == 2) ? 0 : (j + 1);
    float const lm_i = verts[mid][i] - verts[lo][i];
    float const lm_j = verts[mid][j] - verts[lo][j];
    float const lm_k = verts[mid][k] - verts[lo][k];
    float const lh_i = verts[hi][i] - verts[lo][i];
    float const lh_j = verts[hi][j] - verts[lo][j];
    float const lh_k = verts[hi][k] - verts[lo][k];
    float const hm_i = verts[mid][i] - verts[hi][i];
    float const hm_j = verts[mid][j] - verts[hi][j];
    float const hm_k = verts[mid][k] - verts[hi][k];
    float const frac = (lh_i > 0.0f) ? (lm_i / lh_i) : 0.0f;
    float const cut_j = frac * lh_j - lm_j;
    float const cut_k = frac * lh_k - lm_k;
    float const wl0 = cut_j * lm_k - cut_k * lm_j;
    float const wl1 = cut_k * lm_i;
    float const wl2 = cut_j * lm_i;
    float const wu0 = cut_j * hm_k - cut_k * hm_j;
    float const wu1 = cut_k * hm_i;
    float const wu2 = cut_j * hm_i;
    float const area_lo = sqrtf(fmax_ref(wl0 * wl0 + wl1 * wl1 + wl2 * wl2, 0.0f));
    float const area_hi = sqrtf(fmax_ref(wu0 * wu0 + wu1 * wu1 + wu2 * wu2, 0.0f));
    float const rlo_i = verts[lo][i] - P[i];
    float const rhi_i = verts[hi][i] - P[i];
    *out_ii = area_lo * (0.5f * rlo_i * rlo_i + float(2.0 / 3.0) * rlo_i * lm_i + 0.25f * lm_i * lm_i) +
              area_hi * (0.5f * rhi_i * rhi_i + float(2.0 / 3.0) * rhi_i * hm_i + 0.25f * hm_i * hm_i);
    for (int pass = 0; pass < 2; ++pass) {
        float *out = pass == 0 ? out_ij : out_ik;
        int const jk = pass == 0 ? j : k;
        float const cut = pass == 0 ? cut_j : cut_k;
        if (out == nullptr)
            continue;
        float const split = verts[mid][jk] + 0.5f * cut;
        float const lo_split = split - verts[lo][jk];
        float const hi_split = split - verts[hi][jk];
        float const rlo_j = verts[lo][jk] - P[jk];
        float const rhi_j = verts[hi][jk] - P[jk];
        *out = area_lo * (0.5f * rlo_i * rlo_j + float(1.0 / 3.0) * rlo_i * lo_split +
                          float(1.0 / 3.0) * rlo_j * lm_i + 0.25f * lm_i * lo_split) +
               area_hi * (0.5f * rhi_i * rhi_j + float(1.0 / 3.0) * rhi_i * hi_split +
                          float(1.0 / 3.0) * rhi_j * hm_i + 0.25f * hm_i * hi_split);
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

// One node of a level pass: the payload of each child slot (leaf payloads straight from the
// triangle records, internal ones from the child's total written by the previous, deeper level),
// the node's recentred coefficients and slot areas, and the node's total in slot order from the
// first valid slot (the reference's combine_node, gwn_bvh_refit_moment_impl.cuh:17-99).  The whole
// record is written, empty slots and padding as zeros, so the output needs no clearing pass.
template <int Order>
__device__ inline void moment_refit_node(
    std::size_t n, bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, payload<Order> *__restrict__ pending, float4 *__restrict__ records,
    float *__restrict__ slot_area, std::uint32_t *error
) {
    constexpr int RV = moment_record_vec4(Order);
    constexpr int F = moment_field_count(Order);
    float4 *rec = records + n * RV;
    float *fields = reinterpret_cast<float *>(rec + 1);
    std::uint32_t cref[4] = {k_invalid, k_invalid, k_invalid, k_invalid};
    payload<Order> total;
    payload_zero(total);
    bool init = false;
    for (int s = 0; s < 4; ++s) {
        bvh_child const ch = nodes[n].children[s];
        std::uint64_t const ref = ch.reference;
        if (!ref_valid(ref)) {
            for (int f = 0; f < F; ++f)
                fields[f * 4 + s] = 0.0f;
            slot_area[n * 4 + s] = 0.0f;
            continue;
        }
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
    for (int k = 4 + F * 4; k < RV * 4; ++k)
        reinterpret_cast<float *>(rec)[k] = 0.0f;
    rec[0] = make_float4(__uint_as_float(cref[0]), __uint_as_float(cref[1]), __uint_as_float(cref[2]),
                         __uint_as_float(cref[3]));
    pending[n] = total;
}

// One breadth-first level [first, first + count) of one tree; levels run deepest first, so every
// internal child's total is complete (a kernel boundary orders the levels: no atomics, no fences).
template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t tri_count, std::size_t first, std::size_t count, payload<Order> *__restrict__ pending,
    float4 *__restrict__ records, float *__restrict__ slot_area, std::uint32_t *error
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= count)
        return;
    moment_refit_node<Order>(first + i, nodes, node_count, tris, tri_count, pending, records, slot_area, error);
}

// Batched refits: one launch per depth over that depth's nodes in every mesh.  Thread g finds its
// segment (one mesh's node range at this depth) by binary search over the segments' prefixes.
__device__ inline std::uint32_t segment_of(refit_segment const *seg, std::uint32_t count, std::uint64_t g) {
    std::uint32_t lo = 0, hi = count;  // seg[lo].prefix <= g < seg[hi].prefix
    while (hi - lo > 1) {
        std::uint32_t const mid = lo + (hi - lo) / 2;
        if (seg[mid].prefix <= g)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <int Order>
__global__ __launch_bounds__(k_block) void moment_level_batch_kernel(
    moment_refit_mesh const *__restrict__ meshes, refit_segment const *__restrict__ seg, std::uint32_t seg_count,
    std::uint64_t total, std::uint32_t *error
) {
    std::uint64_t const g = static_cast<std::uint64_t>(blockIdx.x) * k_block + threadIdx.x;
    if (g >= total)
        return;
    refit_segment const sg = seg[segment_of(seg, seg_count, g)];
    moment_refit_mesh const m = meshes[sg.mesh];
    std::size_t const n = static_cast<std::size_t>(sg.first_node) + static_cast<std::size_t>(g - sg.prefix);
    if (n >= m.node_count) {
        atomicOr(error, 1u);
        return;
    }
    moment_refit_node<Order>(n, m.nodes, m.node_count, m.tris, m.tri_count, static_cast<payload<Order> *>(m.pending),
                             m.records, m.slot_area, error);
}

template <int Order>
void launch_levels(gwn_bvh_handle const &bvh, gwn_moment_handle &m, void *pending, std::uint32_t *error,
                   hipStream_t stream) {
    std::size_t const N = bvh.node_count;
    auto const &ls = bvh.level_start;
    for (std::size_t d = ls.size() - 1; d-- > 0;) {
        std::size_t const first = ls[d], count = ls[d + 1] - ls[d];
        if (count == 0)
            continue;
        moment_level_kernel<Order><<<grid_for(count, k_block), k_block, 0, stream>>>(
            bvh.nodes.ptr, N, bvh.triangles.ptr, bvh.triangle_count, first, count,
            static_cast<payload<Order> *>(pending), m.records.ptr, m.slot_area.ptr, error);
        GWN_LAUNCH_CHECK();
    }
}

void check_levels(gwn_bvh_handle const &bvh) {
    auto const &ls = bvh.level_start;
    if (ls.size() < 2 || ls.front() != 0u || ls.back() != bvh.node_count)
        throw_internal("bvh.refit.moment: BVH level table does not cover its nodes.");
    for (std::size_t d = 1; d < ls.size(); ++d)
        if (ls[d] < ls[d - 1])
            throw_internal("bvh.refit.moment: BVH level table is not ascending.");
}

std::size_t payload_bytes(int order) {
    return order == 0 ? sizeof(payload<0>) : (order == 1 ? sizeof(payload<1>) : sizeof(payload<2>));
}

void enqueue_levels(gwn_bvh_handle const &bvh, gwn_moment_handle &m, void *pending, std::uint32_t *error,
                    hipStream_t stream) {
    switch (m.order) {
    case 0: launch_levels<0>(bvh, m, pending, error, stream); break;
    case 1: launch_levels<1>(bvh, m, pending, error, stream); break;
    case 2: launch_levels<2>(bvh, m, pending, error, stream); break;
    default: throw_invalid("Taylor order must be 0, 1, or 2.");
    }
}

} // namespace

void refit_moment_prepare(gwn_bvh_handle const &bvh, gwn_moment_handle &m, moment_refit_work &w, hipStream_t stream) {
    std::size_t const N = ref_internal(bvh.root.reference) ? bvh.node_count : 0;
    m.node_count = bvh.node_count;
    m.storage_generation = next_storage_generation();
    if (N == 0)
        return;
    check_levels(bvh);
    m.records.allocate(N * static_cast<std::size_t>(moment_record_vec4(m.order)), stream);
    m.slot_area.allocate(N * 4, stream);
    w.pending.allocate(N * payload_bytes(m.order), stream);
}

void refit_moment_batch(moment_refit_mesh const *meshes, refit_segment const *segments,
                        std::vector<refit_level> const &levels, int order, std::uint32_t *error,
                        hipStream_t stream) {
    for (std::size_t d = levels.size(); d-- > 0;) {
        refit_level const &lv = levels[d];
        if (lv.nodes == 0)
            continue;
        unsigned const grid = grid_for(lv.nodes, k_block);
        refit_segment const *seg = segments + lv.first_segment;
        switch (order) {
        case 0: moment_level_batch_kernel<0><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        case 1: moment_level_batch_kernel<1><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        case 2: moment_level_batch_kernel<2><<<grid, k_block, 0, stream>>>(meshes, seg, lv.segments, lv.nodes, error); break;
        default: throw_invalid("Taylor order must be 0, 1, or 2.");
        }
        GWN_LAUNCH_CHECK();
    }
}

std::size_t moment_payload_bytes(int order) { return payload_bytes(order); }

void refit_moment(gwn_bvh_handle const &bvh, gwn_moment_handle &m, hipStream_t stream) {
    m.node_count = bvh.node_count;
    m.bvh_revision = bvh.revision;
    if (!ref_internal(bvh.root.reference) || bvh.node_count == 0)
        return; // a leaf root has no moment nodes (gwn_bvh.cuh:290-291)
    if (m.order < 0 || m.order > 2)
        throw_invalid("Taylor order must be 0, 1, or 2.");
    check_levels(bvh);
    std::size_t const N = bvh.node_count;
    m.records.allocate(N * static_cast<std::size_t>(moment_record_vec4(m.order)), stream);
    m.slot_area.allocate(N * 4, stream);
    device_buffer<std::uint8_t> pending(N * payload_bytes(m.order), stream);
    device_buffer<std::uint32_t> error(1, stream);
    fill_u32(error.ptr, 0u, 1, stream);
    enqueue_levels(bvh, m, pending.ptr, error.ptr, stream);
    std::uint32_t host_error = 0;
    GWN_HIP(hipMemcpyAsync(&host_error, error.ptr, sizeof(host_error), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host_error != 0u)
        throw_internal("bvh.refit.moment: Bottom-up moment refit did not converge.");
}

} // namespace gwn
