// gwn_winding.hip - exact and Taylor winding-number batches on gfx950.
//
// Reference: include/gwn/detail/gwn_query_winding_impl.cuh:102-443,
// include/gwn/detail/gwn_query_geometry_impl.cuh:16-45, include/gwn/gwn_query.cuh:296-445.
//
// Exact batch: one query per lane; the block streams the leaf-ordered triangle records through
// LDS in 256-record tiles, so every record is read from HBM/L2 once per block and broadcast to
// all 256 lanes.  VALU bound (solid angle: 3 sqrt, 3 div, 1 atan2 per pair).
//
// Taylor batch: one query per lane, queries optionally visited in Hilbert order (results are
// written back in input order).  Per-lane stack of internal node ids; the last internal child
// of a node is entered directly, the others are pushed, leaves are evaluated in slot order --
// the reference's traversal order, so float sums are formed in the same sequence.  The node
// data is one traversal record per node (compact child refs + field-SoA coefficients), so a
// node visit touches 2-3 consecutive 128-byte lines and no separate hierarchy array.
//
// Precision split: traversal and far-field Taylor terms in FP32; a leaf triangle within four edge
// lengths of the query in FP64 on the exact vertices, farther leaf triangles in FP32; every term
// accumulated in FP64.
//
// Products are contracted into FMAs only within one source expression (contract(on), not the
// default across statements), so every kernel that evaluates a term -- per-lane, wave, gradient,
// batched -- rounds it identically and their results agree bit for bit.
#include <mutex>

#include "gwn_core.hpp"
#include "../../include/gwn/detail/gwn_device_winding.hpp"

#include <cmath>

#pragma clang fp contract(on)

namespace gwn {
namespace {

using namespace ::gwn::detail::winding;
constexpr std::uint32_t k_invalid = 0xffffffffu;

// ------------------------------------------------------------------------------------------
// Exact batch
// ------------------------------------------------------------------------------------------

constexpr int k_exact_block = 256;

template <int Prec, class Out>
__global__ __launch_bounds__(k_exact_block) void exact_kernel(
    tri_record const *__restrict__ tris, std::size_t nt, float const *__restrict__ qx,
    float const *__restrict__ qy, float const *__restrict__ qz, Out *__restrict__ out, std::size_t nq
) {
    __shared__ tri_record tile[k_exact_block];
    std::size_t const qi = static_cast<std::size_t>(blockIdx.x) * k_exact_block + threadIdx.x;
    bool const active = qi < nq;
    float const x = active ? qx[qi] : 0.0f;
    float const y = active ? qy[qi] : 0.0f;
    float const z = active ? qz[qi] : 0.0f;
    float sum_f = 0.0f;   // plain: the reference's running float sum
    double sum_d = 0.0;   // split / fp64
    for (std::size_t base = 0; base < nt; base += k_exact_block) {
        std::size_t const count = (nt - base) < k_exact_block ? (nt - base) : k_exact_block;
        __syncthreads();
        {
            float const *src = reinterpret_cast<float const *>(tris + base);
            float *dst = reinterpret_cast<float *>(tile);
            std::size_t const words = count * 9;
            for (std::size_t w = threadIdx.x; w < words; w += k_exact_block)
                dst[w] = src[w];
        }
        __syncthreads();
        if constexpr (Prec == GWN_PRECISION_PLAIN) {
            for (std::size_t k = 0; k < count; ++k)
                sum_f += solid_angle_plain(tile[k], x, y, z);
        } else if constexpr (Prec == GWN_PRECISION_FP64) {
            for (std::size_t k = 0; k < count; ++k)
                sum_d += solid_angle_f64(tile[k], x, y, z);
        } else {
            // Per-tile FP32 partial sums of the far triangles, folded into the FP64 total.
            float far = 0.0f;
            for (std::size_t k = 0; k < count; ++k) {
                tri_record const &t = tile[k];
                float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
                float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
                float const ax = t.v0x - x, ay = t.v0y - y, az = t.v0z - z;
                float const e1 = e1x * e1x + e1y * e1y + e1z * e1z;
                float const e2 = e2x * e2x + e2y * e2y + e2z * e2z;
                float const dq = ax * ax + ay * ay + az * az;
                if (!(dq > k_split_near_scale2 * fmaxf(e1, e2)))
                    sum_d += solid_angle_split_near(t, x, y, z);
                else
                    far += solid_angle_far_f32(ax, ay, az, e1x, e1y, e1z, e2x, e2y, e2z);
            }
            sum_d += double(far);
        }
    }
    if (!active)
        return;
    if constexpr (Prec == GWN_PRECISION_PLAIN)
        out[qi] = static_cast<Out>(sum_f / (4.0f * k_pi_f));
    else
        out[qi] = static_cast<Out>(sum_d / (4.0 * k_pi_d));
}

// ------------------------------------------------------------------------------------------
// Taylor batch
// ------------------------------------------------------------------------------------------

constexpr int k_taylor_block = 256;

__device__ inline std::uint32_t spread10(std::uint32_t v) {
    std::uint32_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Position of a cell of the 1024^3 grid along the 3-D Hilbert curve (30 bits), by Skilling's
// transpose method ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004): undo the
// excess work on the axes' bits from the top down, Gray-encode, then interleave the transposed
// bits (axis 0 most significant in each triple).  Consecutive keys are face-adjacent cells, so a
// wave of consecutive queries is more compact than along the Morton curve, whose keys jump.
__device__ inline std::uint32_t hilbert_key10(std::uint32_t x0, std::uint32_t x1, std::uint32_t x2) {
    std::uint32_t X[3] = {x0 & 0x3ffu, x1 & 0x3ffu, x2 & 0x3ffu};
    for (std::uint32_t Q = 1u << 9; Q > 1u; Q >>= 1) {
        std::uint32_t const P = Q - 1u;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                std::uint32_t const t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    std::uint32_t t = 0;
    for (std::uint32_t Q = 1u << 9; Q > 1u; Q >>= 1)
        if (X[2] & Q)
            t ^= Q - 1u;
    X[0] ^= t;
    X[1] ^= t;
    X[2] ^= t;
    return (spread10(X[0]) << 2) | (spread10(X[1]) << 1) | spread10(X[2]);
}

// Sort keys of the query presort: each query's Hilbert position on the 1024^3 grid over the
// BVH's root box (points outside are clamped onto it).
__global__ __launch_bounds__(256) void query_key_kernel(
    float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::size_t n, aabb box, std::uint32_t *__restrict__ keys, std::uint32_t *__restrict__ ids
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n)
        return;
    float const sx = box.max_x > box.min_x ? 1.0f / (box.max_x - box.min_x) : 1.0f;
    float const sy = box.max_y > box.min_y ? 1.0f / (box.max_y - box.min_y) : 1.0f;
    float const sz = box.max_z > box.min_z ? 1.0f / (box.max_z - box.min_z) : 1.0f;
    // NaN coordinates compare false and land in cell 0; order only affects performance.
    float const fx = fminf(fmaxf((qx[i] - box.min_x) * sx, 0.0f), 1.0f) * 1023.0f;
    float const fy = fminf(fmaxf((qy[i] - box.min_y) * sy, 0.0f), 1.0f) * 1023.0f;
    float const fz = fminf(fmaxf((qz[i] - box.min_z) * sz, 0.0f), 1.0f) * 1023.0f;
    keys[i] = hilbert_key10(static_cast<std::uint32_t>(fx), static_cast<std::uint32_t>(fy),
                            static_cast<std::uint32_t>(fz));
    ids[i] = static_cast<std::uint32_t>(i);
}


template <int Order, int Prec, bool Grad, int StackCap>
__global__ __launch_bounds__(k_taylor_block) void taylor_kernel(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node,
    float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::uint32_t const *__restrict__ order, taylor_outputs o, std::size_t n, float beta2
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_taylor_block + threadIdx.x;
    if (i >= n)
        return;
    std::size_t const q = order ? order[i] : i;
    float const x = qx[q], y = qy[q], z = qz[q];
    taylor_sums<Prec, Grad> acc;
    std::uint32_t n_tests = 0, n_accept = 0, n_leaf = 0;
    bool const overflow =
        taylor_traverse<Order, Prec, Grad, StackCap>(records, tris, root_node, x, y, z, beta2, acc, n_tests, n_accept, n_leaf);
    acc.store(overflow, q, o.out, o.gx, o.gy, o.gz);
    if (o.counters != nullptr) {
        o.counters[3 * q + 0] = n_tests;
        o.counters[3 * q + 1] = n_accept;
        o.counters[3 * q + 2] = n_leaf;
    }
}

// Wave-cooperative traversal (defined below; the batched kernel uses it too).
template <int Order, int Prec, bool Grad, int StackCap>
__device__ inline bool taylor_wave_traverse(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node, bool active,
    float x, float y, float z, float beta2, taylor_sums<Prec, Grad> &acc, std::uint32_t &n_tests,
    std::uint32_t &n_accept, std::uint32_t &n_leaf, std::uint32_t *s_node, std::uint64_t *s_mask
);

// Batched persistent traversal: many meshes, one launch.  Work is a list of chunks of `chunk`
// consecutive queries of one mesh; every block takes chunks from a global counter until none is
// left (no wave waits for another), and its lanes evaluate one query each with that mesh's
// records and the per-lane traversal above, so each result equals the single-mesh batch's.
// Mode 0: one query per lane; 1: wave-cooperative; 2: each mesh as its `wave` flag says (a wave
// walks the union of its lanes' paths, which pays where a mesh's queries are dense and costs
// where a few queries are spread over a large tree).  Every mode gives the same results.
template <int Order, int StackCap, int Mode>
__global__ __launch_bounds__(256) void taylor_batch_kernel(
    batch_mesh const *__restrict__ meshes, std::uint32_t mesh_count, std::uint64_t total_chunks, std::uint32_t chunk,
    float beta2, std::uint64_t *counter
) {
    __shared__ std::uint64_t s_chunk;
    __shared__ std::uint32_t s_node[4][Mode != 0 ? StackCap : 1];
    __shared__ std::uint64_t s_mask[4][Mode != 0 ? StackCap : 1];
    int const wave = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk = atomicAdd(reinterpret_cast<unsigned long long *>(counter), 1ull);
        __syncthreads();
        std::uint64_t const c = s_chunk;
        __syncthreads();
        if (c >= total_chunks)
            return;
        // Mesh of this chunk: last mesh whose first_chunk <= c (binary search).
        std::uint32_t lo = 0, hi = mesh_count;
        while (hi - lo > 1) {
            std::uint32_t const mid = (lo + hi) >> 1;
            if (meshes[mid].first_chunk <= c)
                lo = mid;
            else
                hi = mid;
        }
        batch_mesh const m = meshes[lo];
        std::uint64_t const base = (c - m.first_chunk) * chunk;
        bool const leaf_root = ref_leaf(m.root);
        // Whole waves step through the chunk together, so the wave traversal sees uniform meshes.
        for (std::uint32_t t = threadIdx.x; t - (threadIdx.x & 63u) < chunk; t += blockDim.x) {
            std::uint64_t const p = base + t;
            bool const active = t < chunk && p < m.count;
            std::uint64_t const q = active ? (m.order ? m.order[p] : p) : 0;
            float const x = active ? m.qx[q] : 0.0f, y = active ? m.qy[q] : 0.0f, z = active ? m.qz[q] : 0.0f;
            taylor_sums<GWN_PRECISION_PLAIN, false> acc;
            std::uint32_t n_tests = 0, n_accept = 0, n_leaf = 0;
            bool overflow = false;
            if (leaf_root) {
                if (!active)
                    continue;
                std::uint64_t const off = ref_offset(m.root);
                std::uint32_t const cnt = ref_count(m.root);
                for (std::uint32_t k = 0; k < cnt; ++k)
                    acc.leaf(m.tris[off + k], x, y, z);
            } else if (Mode == 1 || (Mode == 2 && m.wave != 0u)) {  // uniform over the block
                if constexpr (Mode != 0) {
                    if (__ballot(active) == 0ull)
                        continue;
                    overflow = taylor_wave_traverse<Order, GWN_PRECISION_PLAIN, false, StackCap>(
                        m.records, m.tris, static_cast<std::uint32_t>(ref_offset(m.root)), active, x, y, z, beta2,
                        acc, n_tests, n_accept, n_leaf, s_node[wave], s_mask[wave]);
                    if (!active)
                        continue;
                }
            } else {
                if (!active)
                    continue;
                overflow = taylor_traverse<Order, GWN_PRECISION_PLAIN, false, StackCap>(
                    m.records, m.tris, static_cast<std::uint32_t>(ref_offset(m.root)), x, y, z, beta2, acc, n_tests,
                    n_accept, n_leaf);
            }
            acc.store(overflow, q, m.out, nullptr, nullptr, nullptr);
        }
    }
}

// Wave-cooperative traversal for spatially sorted queries.  The 64 lanes of a wave walk the union
// of their traversal paths together: the wave-uniform stack holds (node, lane mask) pairs, every
// node record is fetched once per wave from a wave-uniform address, and each lane only tests the
// children of nodes its own traversal reaches.  A lane's descended internal children are pushed
// and entered in the same slot / last-in-first-out order as in the per-lane kernel, so every lane
// forms its sum over exactly the same terms in exactly the same order; the stack depth is bounded
// by internal_stack_bound for the same reason.
// The wave-cooperative traversal of one wave's queries (lanes with `active`); s_node / s_mask are
// this wave's LDS stack.  The wave must hold at least one active lane.  Returns true on overflow.
// `records` are evaluation records for the value alone and traversal records with the gradient,
// as for the per-lane traversal.
template <int Order, int Prec, bool Grad, int StackCap>
__device__ inline bool taylor_wave_traverse(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node, bool active,
    float x, float y, float z, float beta2, taylor_sums<Prec, Grad> &acc, std::uint32_t &n_tests,
    std::uint32_t &n_accept, std::uint32_t &n_leaf, std::uint32_t *s_node, std::uint64_t *s_mask
) {
    constexpr int RV = Grad ? moment_record_vec4(Order) : eval_record_vec4(Order);
    constexpr int B = eval_slot_floats(Order);
    int const lane = threadIdx.x & 63;
    std::uint64_t mask = __ballot(active);
    std::uint64_t const my_bit = 1ull << lane;
    std::uint32_t node = root_node;
    int sp = 0;
    bool overflow = false;  // wave-uniform
    while (node != k_invalid && !overflow) {
        node = __builtin_amdgcn_readfirstlane(node);
        float4 const *rec = records + static_cast<std::size_t>(node) * RV;
        float4 const refs = rec[0];
        float4 const mp = rec[1], px = rec[2], py = rec[3], pz = rec[4];
        float const *fields = reinterpret_cast<float const *>(rec + 1);
        std::uint32_t const cref[4] = {__float_as_uint(refs.x), __float_as_uint(refs.y),
                                       __float_as_uint(refs.z), __float_as_uint(refs.w)};
        float const mpd[4] = {mp.x, mp.y, mp.z, mp.w};
        float const cx[4] = {px.x, px.y, px.z, px.w};
        float const cy[4] = {py.x, py.y, py.z, py.w};
        float const cz[4] = {pz.x, pz.y, pz.z, pz.w};
        bool const in = (mask & my_bit) != 0ull;
        std::uint32_t next = k_invalid;
        std::uint64_t next_mask = 0ull;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            std::uint32_t const c = __builtin_amdgcn_readfirstlane(cref[s]);
            if (c == k_invalid || overflow)
                continue;
            bool descend = false;
            if (in) {
                ++n_tests;
                float const qrx = x - cx[s], qry = y - cy[s], qrz = z - cz[s];
                float const ql2 = squared_norm_unfused(qrx, qry, qrz);
                descend = !(ql2 > 0.0f) || ql2 <= product_unfused(mpd[s], beta2);
                if (!descend) {
                    float g[3];
                    float w;
                    if constexpr (Grad)
                        w = taylor_term<Order, true>(fields, s, qrx, qry, qrz, ql2, g);
                    else
                        w = taylor_value<Order>(reinterpret_cast<float const *>(rec + 5) + s * B, qrx, qry, qrz,
                                                ql2);
                    if (isfinite(w)) {
                        acc.far_term(w, g);
                        ++n_accept;
                    } else {
                        descend = true;
                    }
                }
            }
            std::uint64_t const dmask = __ballot(descend);
            if (dmask == 0ull)
                continue;
            if (c & k_cref_leaf) {
                tri_record const t = tris[c & ~k_cref_leaf];
                if (dmask & my_bit) {
                    ++n_leaf;
                    acc.leaf(t, x, y, z);
                }
                continue;
            }
            if (next != k_invalid) {
                if (sp < StackCap) {
                    s_node[sp] = next;
                    s_mask[sp] = next_mask;
                    ++sp;
                } else {
                    overflow = true;
                }
            }
            next = c;
            next_mask = dmask;
        }
        if (next == k_invalid && sp > 0) {
            --sp;
            next = s_node[sp];
            next_mask = s_mask[sp];
        }
        node = next;
        mask = next_mask;
    }
    return overflow;
}

template <int Order, int Prec, bool Grad, int StackCap>
__global__ __launch_bounds__(k_taylor_block) void taylor_wave_kernel(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node,
    float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::uint32_t const *__restrict__ order, taylor_outputs o, std::size_t n, float beta2
) {
    constexpr int k_waves = k_taylor_block / 64;
    __shared__ std::uint32_t s_node[k_waves][StackCap];
    __shared__ std::uint64_t s_mask[k_waves][StackCap];
    int const wave = threadIdx.x >> 6;
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_taylor_block + threadIdx.x;
    bool const active = i < n;
    std::size_t const q = active ? (order ? order[i] : i) : 0;
    float const x = active ? qx[q] : 0.0f;
    float const y = active ? qy[q] : 0.0f;
    float const z = active ? qz[q] : 0.0f;
    if (__ballot(active) == 0ull)
        return;
    taylor_sums<Prec, Grad> acc;
    std::uint32_t n_tests = 0, n_accept = 0, n_leaf = 0;
    bool const overflow = taylor_wave_traverse<Order, Prec, Grad, StackCap>(
        records, tris, root_node, active, x, y, z, beta2, acc, n_tests, n_accept, n_leaf, s_node[wave], s_mask[wave]);
    if (!active)
        return;
    acc.store(overflow, q, o.out, o.gx, o.gy, o.gz);
    if (o.counters != nullptr) {
        o.counters[3 * q + 0] = n_tests;
        o.counters[3 * q + 1] = n_accept;
        o.counters[3 * q + 2] = n_leaf;
    }
}

template <int Prec, bool Grad>
__global__ __launch_bounds__(256) void leaf_root_kernel(
    tri_record const *__restrict__ tris, std::uint64_t offset, std::uint32_t count,
    float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    taylor_outputs o, std::size_t n
) {
    std::size_t const q = static_cast<std::size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (q >= n)
        return;
    float const x = qx[q], y = qy[q], z = qz[q];
    taylor_sums<Prec, Grad> acc;
    for (std::uint32_t k = 0; k < count; ++k)
        acc.leaf(tris[offset + k], x, y, z);
    acc.store(false, q, o.out, o.gx, o.gy, o.gz);
    if (o.counters != nullptr) {
        o.counters[3 * q + 0] = 0;
        o.counters[3 * q + 1] = 0;
        o.counters[3 * q + 2] = count;
    }
}

// Reciprocal square root and reciprocal in FP64 from the hardware approximations and Newton
// steps (relative error at FP64 rounding level, not correctly rounded).  x > 0 / x != 0.
__device__ inline double rsq_f64_fast(double x) {
    double r = __builtin_amdgcn_rsq(x);
    double const h = 0.5 * x;
    r = r * fma(-h * r, r, 1.5);
    r = r * fma(-h * r, r, 1.5);
    return r;
}
__device__ inline double rcp_f64_fast(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ inline double rsq_fast(double x) { return rsq_f64_fast(x); }
__device__ inline float rsq_fast(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rcp_fast(double x) { return rcp_f64_fast(x); }
__device__ inline float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }

// Derivatives of one triangle's solid angle W(a, b, c) seen from the origin of a, b, c (vertices
// relative to the query), from the unnormalised form W = 2 atan2(a.(b x c), D) with
// D = |a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a| (equal to the Van Oosterom-Strackee quotient
// scaled by |a||b||c|):  dW = 2 (D dnum - num dD) / (num^2 + D^2).  Zero where W is defined as 0.
template <class R>
__device__ inline void solid_angle_vertex_grad(
    R const (&a)[3], R const (&b)[3], R const (&c)[3], R (&da)[3], R (&db)[3], R (&dc)[3]
) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
        da[k] = db[k] = dc[k] = R(0);
    R const aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    R const bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    R const cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
    if (aa == R(0) || bb == R(0) || cc == R(0))
        return;
    R const ia = rsq_fast(aa), ib = rsq_fast(bb), ic = rsq_fast(cc);
    R const la = aa * ia, lb = bb * ib, lc = cc * ic;
    R const bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    R const ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    R const ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    R const num = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    if (num == R(0))
        return;
    R const dab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    R const dac = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    R const dbc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    R const den = la * lb * lc + dab * lc + dac * lb + dbc * la;
    R const sc = R(2) * rcp_fast(num * num + den * den);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R const ha = a[k] * ia, hb = b[k] * ib, hc = c[k] * ic;
        R const dDa = ha * lb * lc + b[k] * lc + c[k] * lb + dbc * ha;
        R const dDb = hb * la * lc + a[k] * lc + c[k] * la + dac * hb;
        R const dDc = hc * la * lb + a[k] * lb + b[k] * la + dab * hc;
        da[k] = sc * (den * bc[k] - num * dDa);
        db[k] = sc * (den * ca[k] - num * dDb);
        dc[k] = sc * (den * ab[k] - num * dDc);
    }
}

// u * dW/d(a, b, c) of one descended triangle into t9 (9 floats) and -(da + db + dc) * u into dq.
// Split like the forward's precision split: within about four edge lengths (where the terms are
// ~1/d^2, cancel between neighbouring triangles and lose ~|a|/d of their FP32 digits) in FP64 from
// exact differences of the float vertices and query; beyond it in FP32.
__device__ inline void adjoint_leaf_terms(
    tri_record const &t, float x, float y, float z, double sud, float (&t9)[9], double (&dq)[3]
) {
    float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    float const ax = t.v0x - x, ay = t.v0y - y, az = t.v0z - z;
    float const le1 = squared_norm_unfused(e1x, e1y, e1z);
    float const le2 = squared_norm_unfused(e2x, e2y, e2z);
    float const dq2 = squared_norm_unfused(ax, ay, az);
    if (!(dq2 > product_unfused(k_split_near_scale2, fmaxf(le1, le2)))) {
        double const qd[3] = {x, y, z};
        double const a[3] = {double(t.v0x) - qd[0], double(t.v0y) - qd[1], double(t.v0z) - qd[2]};
        double const b[3] = {double(t.v1x) - qd[0], double(t.v1y) - qd[1], double(t.v1z) - qd[2]};
        double const c[3] = {double(t.v2x) - qd[0], double(t.v2y) - qd[1], double(t.v2z) - qd[2]};
        double da[3], db[3], dc[3];
        solid_angle_vertex_grad<double>(a, b, c, da, db, dc);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            t9[k] = static_cast<float>(sud * da[k]);
            t9[3 + k] = static_cast<float>(sud * db[k]);
            t9[6 + k] = static_cast<float>(sud * dc[k]);
            dq[k] -= sud * (da[k] + db[k] + dc[k]);
        }
    } else {
        float const a[3] = {ax, ay, az};
        float const b[3] = {t.v1x - x, t.v1y - y, t.v1z - z};
        float const c[3] = {t.v2x - x, t.v2y - y, t.v2z - z};
        float da[3], db[3], dc[3];
        solid_angle_vertex_grad<float>(a, b, c, da, db, dc);
        float const su = static_cast<float>(sud);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            t9[k] = su * da[k];
            t9[3 + k] = su * db[k];
            t9[6 + k] = su * dc[k];
            dq[k] -= sud * double(da[k] + db[k] + dc[k]);
        }
    }
}

// Adds one descended triangle's adjoint to its record's 9 entries and its query part to dq.
__device__ inline void adjoint_leaf(
    tri_record const &t, std::size_t rec, float x, float y, float z, double sud, float *__restrict__ dtri,
    double (&dq)[3]
) {
    float t9[9];
    adjoint_leaf_terms(t, x, y, z, sud, t9, dq);
    float *d = dtri + rec * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        atomicAdd(d + k, t9[k]);
}

// Adjoint of sum_q u_q w_q (orders 0, 1): the per-lane traversal of taylor_kernel with the same
// accept/descend decisions.  An accepted child adds u/(4 pi) (dw/dP, dw/dN, dw/dT) to its slot's
// 12 entries (dw/dP = -dw/dq); a near triangle adds u/(4 pi) dW/d(a, b, c) to its record's 9
// entries; the query gradient goes to dq.  root_node = k_invalid evaluates the leaf root instead.
template <int Order, int StackCap>
__global__ __launch_bounds__(k_taylor_block) void taylor_adjoint_kernel(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node,
    std::uint64_t root_leaf, float const *__restrict__ qx, float const *__restrict__ qy,
    float const *__restrict__ qz, float const *__restrict__ u, std::uint32_t const *__restrict__ order,
    std::size_t n, float beta2, float *__restrict__ dslot, float *__restrict__ dtri, float *__restrict__ dqx,
    float *__restrict__ dqy, float *__restrict__ dqz
) {
    static_assert(Order <= 1, "adjoint orders are 0 and 1");
    constexpr int RV = moment_record_vec4(Order);
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_taylor_block + threadIdx.x;
    if (i >= n)
        return;
    std::size_t const q = order ? order[i] : i;
    float const x = qx[q], y = qy[q], z = qz[q];
    float const su = u[q] * k_inv_4pi_f;
    double const sud = double(u[q]) * (1.0 / (4.0 * k_pi_d));
    double dq[3] = {0.0, 0.0, 0.0};  // FP64 per-query sum of the far and near parts
    if (root_node == k_invalid) {
        std::uint64_t const off = ref_offset(root_leaf);
        std::uint32_t const cnt = ref_count(root_leaf);
        for (std::uint32_t k = 0; k < cnt; ++k)
            adjoint_leaf(tris[off + k], off + k, x, y, z, sud, dtri, dq);
    } else {
        std::uint32_t stack[StackCap];
        int sp = 0;
        bool overflow = false;
        std::uint32_t node = root_node;
        while (node != k_invalid && !overflow) {
            float4 const *rec = records + static_cast<std::size_t>(node) * RV;
            float4 const refs = rec[0];
            float4 const mp = rec[1], px = rec[2], py = rec[3], pz = rec[4];
            float const *fields = reinterpret_cast<float const *>(rec + 1);
            std::uint32_t const cref[4] = {__float_as_uint(refs.x), __float_as_uint(refs.y),
                                           __float_as_uint(refs.z), __float_as_uint(refs.w)};
            float const mpd[4] = {mp.x, mp.y, mp.z, mp.w};
            float const cx[4] = {px.x, px.y, px.z, px.w};
            float const cy[4] = {py.x, py.y, py.z, py.w};
            float const cz[4] = {pz.x, pz.y, pz.z, pz.w};
            std::uint32_t next = k_invalid;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                std::uint32_t const c = cref[s];
                if (c == k_invalid || overflow)
                    continue;
                float const qrx = x - cx[s], qry = y - cy[s], qrz = z - cz[s];
                float const ql2 = squared_norm_unfused(qrx, qry, qrz);
                bool const descend = !(ql2 > 0.0f) || ql2 <= product_unfused(mpd[s], beta2);
                if (!descend) {
                    float g[3];
                    float const w = taylor_term<Order, true>(fields, s, qrx, qry, qrz, ql2, g);
                    if (isfinite(w)) {
                        float const m1 = __builtin_amdgcn_rsqf(ql2);
                        float const m2 = m1 * m1;
                        float const qn[3] = {qrx * m1, qry * m1, qrz * m1};
                        float *d = dslot + (static_cast<std::size_t>(node) * 4 + s) * 12;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            atomicAdd(d + k, -su * g[k]);
                            atomicAdd(d + 3 + k, -su * m2 * qn[k]);
                            dq[k] += double(su * g[k]);
                        }
                        if constexpr (Order >= 1) {
                            float const m3 = m2 * m1;
                            atomicAdd(d + 6, su * m3 * (1.0f - 3.0f * qn[0] * qn[0]));
                            atomicAdd(d + 7, su * m3 * (1.0f - 3.0f * qn[1] * qn[1]));
                            atomicAdd(d + 8, su * m3 * (1.0f - 3.0f * qn[2] * qn[2]));
                            atomicAdd(d + 9, -3.0f * su * m3 * qn[0] * qn[1]);
                            atomicAdd(d + 10, -3.0f * su * m3 * qn[1] * qn[2]);
                            atomicAdd(d + 11, -3.0f * su * m3 * qn[2] * qn[0]);
                        }
                        continue;
                    }
                }
                if (c & k_cref_leaf) {
                    std::uint32_t const r = c & ~k_cref_leaf;
                    adjoint_leaf(tris[r], r, x, y, z, sud, dtri, dq);
                    continue;
                }
                if (next != k_invalid) {
                    if (sp < StackCap)
                        stack[sp++] = next;
                    else
                        overflow = true;
                }
                next = c;
            }
            if (next == k_invalid && sp > 0)
                next = stack[--sp];
            node = next;
        }
        if (overflow)
            dq[0] = dq[1] = dq[2] = __builtin_nan("");
    }
    if (dqx != nullptr) {
        dqx[q] = static_cast<float>(dq[0]);
        dqy[q] = static_cast<float>(dq[1]);
        dqz[q] = static_cast<float>(dq[2]);
    }
}

template <int Order, int Prec, bool Grad, int Cap>
void launch_taylor_cap(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy,
    float const *qz, std::uint32_t const *order, taylor_outputs const &o, std::size_t n, float beta2,
    hipStream_t stream
) {
    std::uint32_t const root = static_cast<std::uint32_t>(ref_offset(bvh.root.reference));
    // The value alone reads the evaluation records; the gradient the traversal records.
    float4 const *records = Grad ? m.records.ptr : m.eval_records.ptr;
    if (order != nullptr) {
        // Hilbert-ordered queries: neighbouring lanes share most of their paths.
        taylor_wave_kernel<Order, Prec, Grad, Cap><<<grid_for(n, k_taylor_block), k_taylor_block, 0, stream>>>(
            records, bvh.triangles.ptr, root, qx, qy, qz, order, o, n, beta2
        );
    } else {
        taylor_kernel<Order, Prec, Grad, Cap><<<grid_for(n, k_taylor_block), k_taylor_block, 0, stream>>>(
            records, bvh.triangles.ptr, root, qx, qy, qz, order, o, n, beta2
        );
    }
    GWN_LAUNCH_CHECK();
}

template <int Order, int Prec, bool Grad>
void launch_taylor(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy,
    float const *qz, std::uint32_t const *order, taylor_outputs const &o, std::size_t n, float beta2,
    hipStream_t stream
) {
    // The stack only ever holds internal_stack_bound entries; the smallest compiled capacity that
    // covers the bound keeps the LDS / register footprint small.
    int const need = static_cast<int>(bvh.internal_stack_bound);
    if (need <= 32)
        launch_taylor_cap<Order, Prec, Grad, 32>(bvh, m, qx, qy, qz, order, o, n, beta2, stream);
    else if (need <= 64)
        launch_taylor_cap<Order, Prec, Grad, 64>(bvh, m, qx, qy, qz, order, o, n, beta2, stream);
    else if (need <= 128)
        launch_taylor_cap<Order, Prec, Grad, 128>(bvh, m, qx, qy, qz, order, o, n, beta2, stream);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
}


template <int Ctrl, int RowMask>
__device__ inline float dpp_f32(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), Ctrl, RowMask, 0xf, false));
}

// The two results of a v_permlane{16,32}_swap, forced into registers before use: summing them
// straight from the builtin's pair, hipcc (ROCm 7.2) emitted `v_add_f32 vA, vA, vA` -- the first
// result twice -- after the swap; the empty asm keeps both operands distinct.
__device__ inline float swap_pair_sum(int x, int y) {
    asm volatile("" : "+v"(x), "+v"(y));
    return __builtin_bit_cast(float, x) + __builtin_bit_cast(float, y);
}

// Wave totals of 12 per-lane values with a transposing butterfly (gfx950): v_permlane32_swap adds
// the upper half-wave of value k + 6 to the lower half of value k (and the reverse), leaving value k
// in lanes 0-31 and k + 6 in lanes 32-63; v_permlane16_swap does the same between the rows of each
// half; DPP then sums within each 16-lane row.  On return every lane of row r (lanes 16r..16r+15)
// holds in out[k] the wave total of value k + 3r.  About 30 VALU operations instead of 12 DPP
// reductions (84).  Every lane must be active.
__device__ inline void wave_sum12(float const (&f)[12], float (&out)[3]) {
    float g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        auto const p = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(int, f[k]),
                                                        __builtin_bit_cast(int, f[k + 6]), false, false);
        g[k] = swap_pair_sum(p[0], p[1]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        auto const p = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(int, g[k]),
                                                        __builtin_bit_cast(int, g[k + 3]), false, false);
        float h = swap_pair_sum(p[0], p[1]);
        h += dpp_f32<0xb1, 0xf>(h);   // quad_perm [1, 0, 3, 2]
        h += dpp_f32<0x4e, 0xf>(h);   // quad_perm [2, 3, 0, 1]
        h += dpp_f32<0x141, 0xf>(h);  // row_half_mirror
        h += dpp_f32<0x140, 0xf>(h);  // row_mirror
        out[k] = h;
    }
}

// Adds the wave's per-lane values f[0..F) (F <= 12) into d[0..F): wave totals from wave_sum12, then
// one atomic instruction whose active lanes 16r + k (k < 3) add total k + 3r, so the F adds of a
// call go to F consecutive floats in one request.  A lone contributing lane goes the same way (its
// values plus zeros are its values exactly): one request instead of F.
template <int F>
__device__ inline void wave_atomic_add(float *__restrict__ d, float const (&f)[F]) {
    static_assert(F <= 12, "at most 12 values per wave sum");
    float v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k)
        v[k] = k < F ? f[k] : 0.0f;
    float out[3];
    wave_sum12(v, out);
    int const lane = threadIdx.x & 63;
    int const k = lane & 15;
    int const idx = k + 3 * (lane >> 4);
    if (k < 3 && idx < F)
        atomicAdd(d + idx, k == 0 ? out[0] : (k == 1 ? out[1] : out[2]));
}

// The adjoint of taylor_adjoint_kernel, wave-cooperative: a wave walks the tree once for its 64
// (Hilbert-ordered) queries with the same per-query accept/descend decisions, and the field and
// vertex adjoints of a child slot or near triangle are summed over the wave and added with one
// atomic instruction over consecutive floats, instead of every query adding its own (the top
// slots are shared by every query, so per-query atomics serialise on them).
template <int Order, int StackCap>
__global__ __launch_bounds__(k_taylor_block) void taylor_adjoint_wave_kernel(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node,
    float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    float const *__restrict__ u, std::uint32_t const *__restrict__ order, std::size_t n, float beta2,
    float *__restrict__ dslot, float *__restrict__ dtri, float *__restrict__ dqx, float *__restrict__ dqy,
    float *__restrict__ dqz
) {
    static_assert(Order <= 1, "adjoint orders are 0 and 1");
    constexpr int RV = moment_record_vec4(Order);
    constexpr int F = Order >= 1 ? 12 : 6;
    constexpr int k_waves = k_taylor_block / 64;
    __shared__ std::uint32_t s_node[k_waves][StackCap];
    __shared__ std::uint64_t s_mask[k_waves][StackCap];
    int const wave = threadIdx.x >> 6;
    int const lane = threadIdx.x & 63;
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_taylor_block + threadIdx.x;
    bool const active = i < n;
    std::size_t const q = active ? (order ? order[i] : i) : 0;
    float const x = active ? qx[q] : 0.0f;
    float const y = active ? qy[q] : 0.0f;
    float const z = active ? qz[q] : 0.0f;
    float const su = active ? u[q] * k_inv_4pi_f : 0.0f;
    double const sud = active ? double(u[q]) * (1.0 / (4.0 * k_pi_d)) : 0.0;
    std::uint64_t mask = __ballot(active);
    if (mask == 0ull)
        return;
    std::uint64_t const my_bit = 1ull << lane;
    double dq[3] = {0.0, 0.0, 0.0};
    std::uint32_t node = root_node;
    int sp = 0;
    bool overflow = false;  // wave-uniform
    while (node != k_invalid && !overflow) {
        node = __builtin_amdgcn_readfirstlane(node);
        float4 const *rec = records + static_cast<std::size_t>(node) * RV;
        float4 const refs = rec[0];
        float4 const mp = rec[1], px = rec[2], py = rec[3], pz = rec[4];
        float const *fields = reinterpret_cast<float const *>(rec + 1);
        std::uint32_t const cref[4] = {__float_as_uint(refs.x), __float_as_uint(refs.y),
                                       __float_as_uint(refs.z), __float_as_uint(refs.w)};
        float const mpd[4] = {mp.x, mp.y, mp.z, mp.w};
        float const cx[4] = {px.x, px.y, px.z, px.w};
        float const cy[4] = {py.x, py.y, py.z, py.w};
        float const cz[4] = {pz.x, pz.y, pz.z, pz.w};
        bool const in = (mask & my_bit) != 0ull;
        std::uint32_t next = k_invalid;
        std::uint64_t next_mask = 0ull;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            std::uint32_t const c = __builtin_amdgcn_readfirstlane(cref[s]);
            if (c == k_invalid || overflow)
                continue;
            bool descend = false, accepted = false;
            float f[F];
#pragma unroll
            for (int k = 0; k < F; ++k)
                f[k] = 0.0f;
            if (in) {
                float const qrx = x - cx[s], qry = y - cy[s], qrz = z - cz[s];
                float const ql2 = squared_norm_unfused(qrx, qry, qrz);
                descend = !(ql2 > 0.0f) || ql2 <= product_unfused(mpd[s], beta2);
                if (!descend) {
                    float g[3];
                    float const w = taylor_term<Order, true>(fields, s, qrx, qry, qrz, ql2, g);
                    if (isfinite(w)) {
                        accepted = true;
                        float const m1 = __builtin_amdgcn_rsqf(ql2);
                        float const m2 = m1 * m1;
                        float const qn[3] = {qrx * m1, qry * m1, qrz * m1};
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            f[k] = -su * g[k];
                            f[3 + k] = -su * m2 * qn[k];
                            dq[k] += double(su * g[k]);
                        }
                        if constexpr (Order >= 1) {
                            float const m3 = m2 * m1;
                            f[6] = su * m3 * (1.0f - 3.0f * qn[0] * qn[0]);
                            f[7] = su * m3 * (1.0f - 3.0f * qn[1] * qn[1]);
                            f[8] = su * m3 * (1.0f - 3.0f * qn[2] * qn[2]);
                            f[9] = -3.0f * su * m3 * qn[0] * qn[1];
                            f[10] = -3.0f * su * m3 * qn[1] * qn[2];
                            f[11] = -3.0f * su * m3 * qn[2] * qn[0];
                        }
                    } else {
                        descend = true;
                    }
                }
            }
            std::uint64_t const amask = __ballot(accepted);
            if (amask != 0ull)
                wave_atomic_add<F>(dslot + (static_cast<std::size_t>(node) * 4 + s) * 12, f);
            std::uint64_t const dmask = __ballot(descend);
            if (dmask == 0ull)
                continue;
            if (c & k_cref_leaf) {
                std::uint32_t const r = c & ~k_cref_leaf;
                float t9[9];
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    t9[k] = 0.0f;
                if (dmask & my_bit)
                    adjoint_leaf_terms(tris[r], x, y, z, sud, t9, dq);
                wave_atomic_add<9>(dtri + static_cast<std::size_t>(r) * 9, t9);
                continue;
            }
            if (next != k_invalid) {
                if (sp < StackCap) {
                    s_node[wave][sp] = next;
                    s_mask[wave][sp] = next_mask;
                    ++sp;
                } else {
                    overflow = true;
                }
            }
            next = c;
            next_mask = dmask;
        }
        if (next == k_invalid && sp > 0) {
            --sp;
            next = s_node[wave][sp];
            next_mask = s_mask[wave][sp];
        }
        node = next;
        mask = next_mask;
    }
    if (!active)
        return;
    if (overflow)
        dq[0] = dq[1] = dq[2] = __builtin_nan("");
    if (dqx != nullptr) {
        dqx[q] = static_cast<float>(dq[0]);
        dqy[q] = static_cast<float>(dq[1]);
        dqz[q] = static_cast<float>(dq[2]);
    }
}

template <int Order>
void launch_taylor_adjoint(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy, float const *qz,
    float const *u, std::uint32_t const *order, std::size_t n, float beta2, float *dslot, float *dtri,
    float *dqx, float *dqy, float *dqz, hipStream_t stream
) {
    bool const leaf_root = ref_leaf(bvh.root.reference);
    std::uint32_t const root = leaf_root ? k_invalid : static_cast<std::uint32_t>(ref_offset(bvh.root.reference));
    int const need = leaf_root ? 0 : static_cast<int>(bvh.internal_stack_bound);
    float4 const *records = leaf_root ? nullptr : m.records.ptr;
    // Hilbert-ordered queries (large batches) take the wave-cooperative form.
#define GWN_ADJ(C)                                                                                           \
    do {                                                                                                     \
        if (order != nullptr && !leaf_root)                                                                  \
            taylor_adjoint_wave_kernel<Order, C><<<grid_for(n, k_taylor_block), k_taylor_block, 0, stream>>>( \
                records, bvh.triangles.ptr, root, qx, qy, qz, u, order, n, beta2, dslot, dtri, dqx, dqy, dqz);  \
        else                                                                                                 \
            taylor_adjoint_kernel<Order, C><<<grid_for(n, k_taylor_block), k_taylor_block, 0, stream>>>(    \
                records, bvh.triangles.ptr, root, bvh.root.reference, qx, qy, qz, u, order, n, beta2, dslot,   \
                dtri, dqx, dqy, dqz);                                                                        \
    } while (0)
    if (need <= 32)
        GWN_ADJ(32);
    else if (need <= 64)
        GWN_ADJ(64);
    else if (need <= 128)
        GWN_ADJ(128);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
#undef GWN_ADJ
    GWN_LAUNCH_CHECK();
}

template <int Order, int Mode>
void launch_batch_cap(batch_mesh const *meshes, std::uint32_t mesh_count, std::uint64_t total_chunks,
                      std::uint32_t chunk, float beta2, std::uint64_t *counter, int cap, int block, int grid,
                      hipStream_t stream) {
#define GWN_BATCH(C)                                                                                      \
    taylor_batch_kernel<Order, C, Mode><<<grid, block, 0, stream>>>(meshes, mesh_count, total_chunks, chunk, beta2, counter)
    if (cap <= 32)
        GWN_BATCH(32);
    else if (cap <= 64)
        GWN_BATCH(64);
    else if (cap <= 128)
        GWN_BATCH(128);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
#undef GWN_BATCH
    GWN_LAUNCH_CHECK();
}

} // namespace

void taylor_batch_launch(
    batch_mesh const *meshes, std::uint32_t mesh_count, std::uint64_t total_chunks, std::uint32_t chunk,
    float accuracy_scale, std::uint64_t *counter, int order, int cap, int block, int grid, int traversal,
    hipStream_t stream
) {
    float const beta2 = accuracy_scale * accuracy_scale;
    if (block <= 0 || block > 256 || (block & 63) != 0)
        throw_invalid("Batched query block size must be 64, 128, 192 or 256.");
    if (traversal < 0 || traversal > 2)
        throw_invalid("Batched query traversal must be 0 (per lane), 1 (wave) or 2 (per mesh).");
#define GWN_BATCH_ORDER(O)                                                                                      \
    (traversal == 0   ? launch_batch_cap<O, 0>(meshes, mesh_count, total_chunks, chunk, beta2, counter, cap, block, \
                                               grid, stream)                                                     \
     : traversal == 1 ? launch_batch_cap<O, 1>(meshes, mesh_count, total_chunks, chunk, beta2, counter, cap, block, \
                                               grid, stream)                                                     \
                      : launch_batch_cap<O, 2>(meshes, mesh_count, total_chunks, chunk, beta2, counter, cap, block, \
                                               grid, stream))
    switch (order) {
    case 0: GWN_BATCH_ORDER(0); break;
    case 1: GWN_BATCH_ORDER(1); break;
    default: GWN_BATCH_ORDER(2); break;
    }
#undef GWN_BATCH_ORDER
}

void taylor_adjoint_traverse(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy, float const *qz,
    float const *u, std::uint32_t const *order, std::size_t n, float accuracy_scale, float *dslot, float *dtri,
    float *dqx, float *dqy, float *dqz, hipStream_t stream
) {
    float const beta2 = accuracy_scale * accuracy_scale;
    if (m.order == 0)
        launch_taylor_adjoint<0>(bvh, m, qx, qy, qz, u, order, n, beta2, dslot, dtri, dqx, dqy, dqz, stream);
    else if (m.order == 1)
        launch_taylor_adjoint<1>(bvh, m, qx, qy, qz, u, order, n, beta2, dslot, dtri, dqx, dqy, dqz, stream);
    else
        throw_invalid("The winding adjoint supports Taylor orders 0 and 1.");
}

void winding_exact(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, void *out,
    std::size_t n, int precision, query_launch const &, hipStream_t stream
) {
    unsigned const grid = grid_for(n, k_exact_block);
    switch (precision) {
    case -1: // reference-form float output
        exact_kernel<GWN_PRECISION_PLAIN, float><<<grid, k_exact_block, 0, stream>>>(
            bvh.triangles.ptr, bvh.triangle_count, qx, qy, qz, static_cast<float *>(out), n);
        break;
    case GWN_PRECISION_PLAIN:
        exact_kernel<GWN_PRECISION_PLAIN, double><<<grid, k_exact_block, 0, stream>>>(
            bvh.triangles.ptr, bvh.triangle_count, qx, qy, qz, static_cast<double *>(out), n);
        break;
    case GWN_PRECISION_SPLIT:
        exact_kernel<GWN_PRECISION_SPLIT, double><<<grid, k_exact_block, 0, stream>>>(
            bvh.triangles.ptr, bvh.triangle_count, qx, qy, qz, static_cast<double *>(out), n);
        break;
    case GWN_PRECISION_FP64:
        exact_kernel<GWN_PRECISION_FP64, double><<<grid, k_exact_block, 0, stream>>>(
            bvh.triangles.ptr, bvh.triangle_count, qx, qy, qz, static_cast<double *>(out), n);
        break;
    default: throw_invalid("Unknown winding precision mode.");
    }
    GWN_LAUNCH_CHECK();
}

namespace {

// Device words the Hilbert presort of n queries needs besides the order itself: keys, their
// alternate buffer, the alternate id buffer and the sort's own workspace.
std::size_t query_order_workspace_words(std::size_t n) {
    return 3 * n + radix_sort_workspace_words(n, 0, 30);
}

// Hilbert presort with caller-given scratch (query_order_workspace_words(n) words): allocates
// nothing.
void query_order_ws(
    aabb const &box, float const *qx, float const *qy, float const *qz,
    std::uint32_t *order, std::size_t n, std::uint32_t *ws, hipStream_t stream
) {
    std::uint32_t *const keys = ws;
    std::uint32_t *const keys_alt = ws + n;
    std::uint32_t *const ids_alt = ws + 2 * n;
    query_key_kernel<<<grid_for(n, 256), 256, 0, stream>>>(qx, qy, qz, n, box, keys, order);
    GWN_LAUNCH_CHECK();
    std::uint32_t *sk = nullptr;
    std::uint32_t *sorted = nullptr;
    radix_sort_pairs_u32_ws(keys, keys_alt, order, ids_alt, n, 0, 30, ws + 3 * n, &sk, &sorted, stream);
    if (sorted != order)
        GWN_HIP(hipMemcpyAsync(order, sorted, n * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, stream));
}

} // namespace

void query_order(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz,
    std::uint32_t *order, std::size_t n, hipStream_t stream
) {
    query_order_box(bvh.root.bounds, qx, qy, qz, order, n, stream);
}

void query_order_box(
    aabb const &box, float const *qx, float const *qy, float const *qz, std::uint32_t *order, std::size_t n,
    hipStream_t stream
) {
    if (n == 0)
        return;
    device_buffer<std::uint32_t> ws(query_order_workspace_words(n), stream);
    query_order_ws(box, qx, qy, qz, order, n, ws.ptr, stream);
}

void winding_taylor(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy,
    float const *qz, float *out, float *gx, float *gy, float *gz, std::size_t n, float accuracy_scale,
    int precision, query_launch const &cfg, std::uint32_t const *given_order, std::uint32_t *counters,
    hipStream_t stream
) {
    if (n == 0)
        return;
    if (precision != GWN_PRECISION_PLAIN && precision != GWN_PRECISION_SPLIT)
        throw_invalid("Taylor winding supports the plain and split precision modes.");
    bool const grad = gx != nullptr;
    taylor_outputs const o{out, gx, gy, gz, counters};
    if (ref_leaf(bvh.root.reference)) {
        unsigned const grid = grid_for(n, 256);
        std::uint64_t const off = ref_offset(bvh.root.reference);
        std::uint32_t const cnt = ref_count(bvh.root.reference);
#define GWN_LEAF_ROOT(P, G) leaf_root_kernel<P, G><<<grid, 256, 0, stream>>>(bvh.triangles.ptr, off, cnt, qx, qy, qz, o, n)
        if (precision == GWN_PRECISION_PLAIN)
            grad ? GWN_LEAF_ROOT(GWN_PRECISION_PLAIN, true) : GWN_LEAF_ROOT(GWN_PRECISION_PLAIN, false);
        else
            grad ? GWN_LEAF_ROOT(GWN_PRECISION_SPLIT, true) : GWN_LEAF_ROOT(GWN_PRECISION_SPLIT, false);
#undef GWN_LEAF_ROOT
        GWN_LAUNCH_CHECK();
        return;
    }
    bool const sort = given_order == nullptr &&
                      (cfg.sort_queries > 0 || (cfg.sort_queries < 0 && n >= (1u << 16)));
    device_buffer<std::uint32_t> own_order;
    std::uint32_t const *order = given_order;
    if (sort) {
        own_order.allocate(n, stream);
        query_order(bvh, qx, qy, qz, own_order.ptr, n, stream);
        order = own_order.ptr;
    }
    float const beta2 = accuracy_scale * accuracy_scale;
#define GWN_TAYLOR_CASE(O, P, G) launch_taylor<O, P, G>(bvh, m, qx, qy, qz, order, o, n, beta2, stream)
#define GWN_TAYLOR_ORDERS(P, G)                                                                     \
    switch (m.order) {                                                                              \
    case 0: GWN_TAYLOR_CASE(0, P, G); break;                                                        \
    case 1: GWN_TAYLOR_CASE(1, P, G); break;                                                        \
    default: GWN_TAYLOR_CASE(2, P, G); break;                                                       \
    }
    if (precision == GWN_PRECISION_PLAIN) {
        if (grad) {
            GWN_TAYLOR_ORDERS(GWN_PRECISION_PLAIN, true)
        } else {
            GWN_TAYLOR_ORDERS(GWN_PRECISION_PLAIN, false)
        }
    } else {
        if (grad) {
            GWN_TAYLOR_ORDERS(GWN_PRECISION_SPLIT, true)
        } else {
            GWN_TAYLOR_ORDERS(GWN_PRECISION_SPLIT, false)
        }
    }
#undef GWN_TAYLOR_ORDERS
#undef GWN_TAYLOR_CASE
}

// ---------------------------------------------------------------------------------------------
// Allocation-free Taylor batch.  Its only scratch is the Hilbert presort's (the visiting order,
// keys, alternate buffers and the sort's digit table), which lives in one static device array per
// GPU, sized for k_noalloc_max_queries queries.  Host code never passes the array's name to a
// kernel: that name is a host-side stand-in.  The array's device address is looked up once per
// device with hipGetSymbolAddress (with that device current) and that pointer is used instead.
// Every call on one GPU shares the array and leaves the last call's contents in it; the sort
// clears its digit bases itself, and everything else is fully overwritten before it is read.
// Calls that overlap on one GPU (different streams, or host threads) race on the array.
// ---------------------------------------------------------------------------------------------

namespace {

constexpr std::size_t k_noalloc_order_words = k_noalloc_max_queries;
constexpr std::size_t k_noalloc_sort_words =
    3 * k_noalloc_max_queries + 4 * 256 + (k_noalloc_max_queries + 4095) / 4096 * 256;
constexpr std::size_t k_noalloc_scratch_words = k_noalloc_order_words + k_noalloc_sort_words;

__device__ __align__(256) std::uint32_t scratch[k_noalloc_scratch_words];

constexpr int k_max_devices = 64;
std::mutex g_scratch_mutex;
std::uint32_t *g_scratch_ptr[k_max_devices] = {};

// The calling thread's current device's copy of `scratch` (resolved once per device).
std::uint32_t *device_scratch() {
    int dev = 0;
    GWN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= k_max_devices)
        throw_internal("noalloc Taylor batch: device ordinal out of range.");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *ptr = nullptr;
        GWN_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr)
            throw_internal("noalloc Taylor batch: scratch symbol has no device address.");
        g_scratch_ptr[dev] = static_cast<std::uint32_t *>(ptr);
    }
    return g_scratch_ptr[dev];
}

} // namespace

void winding_taylor_noalloc(
    gwn_bvh_handle const &bvh, gwn_moment_handle const &m, float const *qx, float const *qy,
    float const *qz, float *out, std::size_t n, float accuracy_scale, int precision, query_launch const &cfg,
    hipStream_t stream
) {
    if (n > k_noalloc_max_queries)
        throw_invalid("noalloc Taylor batch: query count exceeds the static scratch capacity (1048576 queries).");
    static_assert(k_noalloc_sort_words >= 3 * k_noalloc_max_queries + 4 * 256 + (k_noalloc_max_queries + 4095) / 4096 * 256,
                  "presort scratch must cover query_order_workspace_words(k_noalloc_max_queries)");
    if (n == 0)
        return;
    bool const sort = !ref_leaf(bvh.root.reference) &&
                      (cfg.sort_queries > 0 || (cfg.sort_queries < 0 && n >= (1u << 16)));
    if (!sort) {
        winding_taylor(bvh, m, qx, qy, qz, out, nullptr, nullptr, nullptr, n, accuracy_scale, precision, cfg,
                       nullptr, nullptr, stream);
        return;
    }
    if (query_order_workspace_words(n) > k_noalloc_sort_words)
        throw_internal("noalloc Taylor batch: presort scratch too small.");
    std::uint32_t *const base = device_scratch();
    std::uint32_t *const order = base;
    query_order_ws(bvh.root.bounds, qx, qy, qz, order, n, base + k_noalloc_order_words, stream);
    winding_taylor(bvh, m, qx, qy, qz, out, nullptr, nullptr, nullptr, n, accuracy_scale, precision, cfg, order,
                   nullptr, stream);
}

} // namespace gwn
