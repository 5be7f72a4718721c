// gwn_distance.hip - unsigned point-to-mesh distance batch on gfx950.
//
// Reference: include/gwn/detail/gwn_query_distance_impl.cuh:20-192,
// include/gwn/detail/gwn_query_geometry_impl.cuh:46-114, include/gwn/gwn_query.cuh:109-180.
//
// Near-first culled traversal of the canonical BVH4 (one lane per point in distance_kernel, four
// in distance_quad_kernel, the default, with the same results): the culling band
// (clamped at zero) is the initial squared bound and the result when nothing is closer; children
// whose box is no closer than the current best are skipped; the survivors are ordered by box
// distance with a five-comparator network; leaves of a node are evaluated before its internal
// children are queued, the nearest internal child is entered directly and the others are pushed
// far-to-near.  A query that reaches distance zero stops (zero is the global lower bound).
// Queries may be visited through a Hilbert-order permutation so that a wave's lanes walk nearby
// paths; results land at the input positions.  With sort_queries = 2 the Hilbert-ordered batch
// uses a wave-cooperative traversal instead (distance_wave_traverse).
//
// Arithmetic is unfused and the square root correctly rounded, so every result equals the CPU
// oracle's replay bit for bit.
#include "gwn_core.hpp"
#include "../../include/gwn/detail/gwn_device_distance.hpp"

#pragma clang fp contract(off)

namespace gwn {
namespace {

using namespace ::gwn::detail::distance;
constexpr int k_block = 256;


// Wave-cooperative form of distance_traverse (sort_queries = 2).  The wave walks one node at a
// time (node records are wave-uniform loads); every lane computes its own box distances and prunes
// by its own best, so a lane's result is its minimum over the triangles it reaches.  A child is
// entered when any lane still needs it; leaves of a node are evaluated before its internal
// children, children in the order of the first needing lane's box distances (near first).
// Pruning compares FP32 box and triangle distances, and a box's distance can round above that of
// a triangle inside it, so which of two equally near triangles survives depends on visiting
// order: results can differ from the reference-order traversal in the last bits.  `s_node` is
// this wave's LDS stack.  Returns the lane's status (k_dist_overflow for every lane when the
// shared stack overflows).
template <int StackCap>
__device__ inline int distance_wave_traverse(
    distance_scene const &sc, bool active, float x, float y, float z, float &best, std::uint32_t *best_rec,
    std::uint32_t *s_node
) {
    int status = k_dist_done;
    if (ref_leaf(sc.root_ref)) {
        if (active && !visit_leaf(sc.tris, sc.nt, sc.root_ref, x, y, z, best, best_rec))
            status = k_dist_zero;
        return status;
    }
    if (!ref_internal(sc.root_ref))
        return status;
    constexpr std::uint32_t k_none = 0xffffffffu;
    int sp = 0;
    bool overflow = false;  // wave-uniform
    std::uint32_t node = static_cast<std::uint32_t>(ref_offset(sc.root_ref));
    while (node != k_none && !overflow) {
        node = __builtin_amdgcn_readfirstlane(node);
        std::uint32_t next = k_none;
        if (node < sc.node_count) {
            bvh_node4 const &nd = sc.nodes[node];
            float cd[4], key[4];
            std::uint64_t cr[4], need[4];
            int idx[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bvh_child const ch = nd.children[s];
                cr[s] = ch.reference;
                cd[s] = box_distance2(x, y, z, ch.bounds);
                bool const want = active && status == k_dist_done && ref_valid(ch.reference) && cd[s] < best;
                need[s] = __ballot(want);
                key[s] = INFINITY;
                if (need[s] != 0ull) {
                    int const first = __builtin_ctzll(need[s]);
                    key[s] = __builtin_bit_cast(
                        float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, cd[s]), first));
                }
                idx[s] = s;
            }
            sort4(key, idx);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                int const s = idx[k];
                if (need[s] == 0ull || !ref_leaf(cr[s]))
                    continue;
                if (active && status == k_dist_done && cd[s] < best &&
                    !visit_leaf(sc.tris, sc.nt, cr[s], x, y, z, best, best_rec))
                    status = k_dist_zero;
            }
#pragma unroll
            for (int k = 3; k >= 0; --k) {
                int const s = idx[k];
                if (need[s] == 0ull || !ref_internal(cr[s]) || overflow)
                    continue;
                // Re-check after the leaves: some lanes may no longer need this child.
                if (__ballot(active && status == k_dist_done && cd[s] < best) == 0ull)
                    continue;
                if (next != k_none) {
                    if (sp < StackCap)
                        s_node[sp++] = next;
                    else
                        overflow = true;
                }
                next = static_cast<std::uint32_t>(ref_offset(cr[s]));
            }
        }
        if (next == k_none && sp > 0)
            next = s_node[--sp];
        node = next;
    }
    return overflow ? k_dist_overflow : status;
}

template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_wave_kernel(
    distance_scene sc, float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::uint32_t const *__restrict__ order, float *__restrict__ out, std::size_t n, float culling_band,
    std::uint32_t *__restrict__ nearest
) {
    __shared__ std::uint32_t s_node[k_block / 64][StackCap];
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    bool const active = i < n;
    if (__ballot(active) == 0ull)
        return;
    std::size_t const q = active ? order[i] : 0;
    float const x = active ? qx[q] : 0.0f, y = active ? qy[q] : 0.0f, z = active ? qz[q] : 0.0f;
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    std::uint32_t rec = k_nearest_none;
    int const st = distance_wave_traverse<StackCap>(sc, active, x, y, z, best, nearest ? &rec : nullptr,
                                                    s_node[threadIdx.x >> 6]);
    if (active) {
        out[q] = st == k_dist_zero ? 0.0f : (st == k_dist_overflow ? __builtin_nanf("") : sqrtf(best));
        if (nearest != nullptr)
            nearest[q] = nearest_code(st, rec);
    }
}

// Every lane of a quad (lanes 4k..4k+3) receives lane S's value.
template <int S>
__device__ inline std::uint32_t quad_bcast(std::uint32_t v) {
    return static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int>(v), S * 0x55, 0xf, 0xf, false));  // quad_perm [S,S,S,S]
}

// Quad form of distance_traverse: four lanes per point, lane j of the quad owning child slot j of
// each node.  Lane j computes its child's box distance and, for a leaf child nearer than the best
// at node entry, its triangles' minimum and first record attaining it (a triangle's distance does
// not depend on the best found before it).  The quad shares the four results, and every lane
// replays distance_traverse's steps on them: compaction of the children nearer than the best,
// the same five-comparator sort, leaves near-first each against the running best (a leaf's
// minimum replaces the best when strictly smaller, as visit_leaf's strict per-triangle updates
// do), internal children pushed far-to-near.  Distances, records and statuses are therefore those
// of distance_traverse bit for bit; the speculative leaf work only costs lanes that would
// otherwise idle.  The stack is the quad's column of an LDS array (entry e at stack[e * stride]),
// written with the same value by all four lanes.
template <int StackCap>
__device__ inline int distance_quad_traverse(
    distance_scene const &sc, int j, float x, float y, float z, float &best, std::uint32_t &best_rec,
    std::uint32_t *stack, int stride
) {
    int sp = 0;
    int status = k_dist_done;
    std::uint64_t reference = sc.root_ref;
    while (reference != 0 && status == k_dist_done) {
        std::uint64_t next = 0;
        if (ref_leaf(reference)) {
            if (!visit_leaf(sc.tris, sc.nt, reference, x, y, z, best, &best_rec))
                status = k_dist_zero;
        } else if (ref_internal(reference) && ref_offset(reference) < sc.node_count) {
            bvh_child const ch = sc.nodes[ref_offset(reference)].children[j];
            float const d2 = box_distance2(x, y, z, ch.bounds);
            float lmin = INFINITY;
            std::uint32_t lrec = k_nearest_none;
            if (ref_valid(ch.reference) && d2 < best && ref_leaf(ch.reference))
                visit_leaf(sc.tris, sc.nt, ch.reference, x, y, z, lmin, &lrec);
            std::uint32_t const d_bits = __float_as_uint(d2), m_bits = __float_as_uint(lmin);
            std::uint32_t const r_lo = static_cast<std::uint32_t>(ch.reference);
            std::uint32_t const r_hi = static_cast<std::uint32_t>(ch.reference >> 32);
            float sd[4], sm[4];
            std::uint64_t sr[4];
            std::uint32_t sq[4];
#define GWN_QUAD_TAKE(S)                                                                               \
    sd[S] = __uint_as_float(quad_bcast<S>(d_bits));                                                     \
    sm[S] = __uint_as_float(quad_bcast<S>(m_bits));                                                     \
    sr[S] = (static_cast<std::uint64_t>(quad_bcast<S>(r_hi)) << 32) | quad_bcast<S>(r_lo);              \
    sq[S] = quad_bcast<S>(lrec)
            GWN_QUAD_TAKE(0);
            GWN_QUAD_TAKE(1);
            GWN_QUAD_TAKE(2);
            GWN_QUAD_TAKE(3);
#undef GWN_QUAD_TAKE
            // Compaction of the surviving slots in slot order, as selects on each slot's prefix
            // position (no dynamically indexed arrays, which would live in scratch memory).
            float cd[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, cm[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
            std::uint64_t cr[4] = {0, 0, 0, 0};
            std::uint32_t cq[4] = {k_nearest_none, k_nearest_none, k_nearest_none, k_nearest_none};
            int m = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bool const pass = ref_valid(sr[s]) && !(sd[s] >= best);
#pragma unroll
                for (int t = 0; t <= s; ++t) {
                    if (pass && m == t) {
                        cd[t] = sd[s];
                        cr[t] = sr[s];
                        cm[t] = sm[s];
                        cq[t] = sq[s];
                    }
                }
                m += pass ? 1 : 0;
            }
            auto cswap = [&](int a, int b) {
                if (!(cd[a] > cd[b]))
                    return;
                float const td = cd[a];
                cd[a] = cd[b];
                cd[b] = td;
                std::uint64_t const tr = cr[a];
                cr[a] = cr[b];
                cr[b] = tr;
                float const tm = cm[a];
                cm[a] = cm[b];
                cm[b] = tm;
                std::uint32_t const tq = cq[a];
                cq[a] = cq[b];
                cq[b] = tq;
            };
            cswap(0, 1);
            cswap(2, 3);
            cswap(0, 2);
            cswap(1, 3);
            cswap(1, 2);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (status != k_dist_done || s >= m || !ref_leaf(cr[s]) || !(cd[s] < best))
                    continue;
                if (cm[s] < best) {
                    best = cm[s];
                    best_rec = cq[s];
                }
                if (!(best > 0.0f))
                    status = k_dist_zero;
            }
#pragma unroll
            for (int s = 3; s >= 0; --s) {
                if (status != k_dist_done || s >= m || !ref_internal(cr[s]) || cd[s] >= best)
                    continue;
                if (next != 0) {
                    if (sp < StackCap)
                        stack[(sp++) * stride] = static_cast<std::uint32_t>(ref_offset(next));
                    else
                        status = k_dist_overflow;
                }
                next = cr[s];
            }
        }
        if (next == 0 && sp > 0)
            next = encode_internal(stack[(--sp) * stride]);
        reference = next;
    }
    return status;
}

// Four lanes per point (distance_quad_traverse); the grid covers 4 n lanes and whole quads exit
// together.
template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_quad_kernel(
    distance_scene sc, float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::uint32_t const *__restrict__ order, float *__restrict__ out, std::size_t n, float culling_band,
    std::uint32_t *__restrict__ nearest
) {
    constexpr int k_quads = k_block / 4;
    __shared__ std::uint32_t s_stack[StackCap * k_quads];  // entry e of quad k at [e * k_quads + k]
    std::size_t const i = (static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x) >> 2;
    if (i >= n)
        return;
    int const j = static_cast<int>(threadIdx.x & 3);
    std::size_t const q = order ? order[i] : i;
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    std::uint32_t rec = k_nearest_none;
    int const st = distance_quad_traverse<StackCap>(sc, j, qx[q], qy[q], qz[q], best, rec,
                                                    s_stack + (threadIdx.x >> 2), k_quads);
    if (j == 0) {
        out[q] = st == k_dist_zero ? 0.0f : (st == k_dist_overflow ? __builtin_nanf("") : sqrtf(best));
        if (nearest != nullptr)
            nearest[q] = nearest_code(st, rec);
    }
}

template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_kernel(
    distance_scene sc, float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::uint32_t const *__restrict__ order, float *__restrict__ out, std::size_t n, float culling_band,
    std::uint32_t *__restrict__ nearest
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    std::size_t const q = order ? order[i] : i;
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    std::uint32_t rec = k_nearest_none;
    int const st = distance_traverse<StackCap>(sc, qx[q], qy[q], qz[q], best, nearest ? &rec : nullptr);
    out[q] = st == k_dist_zero ? 0.0f : (st == k_dist_overflow ? __builtin_nanf("") : sqrtf(best));
    if (nearest != nullptr)
        nearest[q] = nearest_code(st, rec);
}

// The one-lane traversal in input order, writing only its work counters (box tests, triangle
// tests) per query.
template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_work_kernel(
    distance_scene sc, float const *__restrict__ qx, float const *__restrict__ qy, float const *__restrict__ qz,
    std::size_t n, float culling_band, std::uint32_t *__restrict__ counters
) {
    std::size_t const q = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (q >= n)
        return;
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    traversal_work w;
    (void)distance_traverse<StackCap>(sc, qx[q], qy[q], qz[q], best, nullptr, &w);
    counters[2 * q] = w.boxes;
    counters[2 * q + 1] = w.tris;
}

// Barycentric coordinates (of b and c) of the closest point of triangle (a, b, c) to p: the
// Voronoi-region cases of point_triangle_distance2.
__device__ inline void closest_barycentrics(
    float const (&p)[3], float const (&a)[3], float const (&b)[3], float const (&c)[3], float &v, float &w
) {
    float const ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    float const ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float const ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    float const d1 = dot3(ab[0], ab[1], ab[2], ap[0], ap[1], ap[2]);
    float const d2 = dot3(ac[0], ac[1], ac[2], ap[0], ap[1], ap[2]);
    v = w = 0.0f;
    if (d1 <= 0.0f && d2 <= 0.0f)
        return;
    float const bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    float const d3 = dot3(ab[0], ab[1], ab[2], bp[0], bp[1], bp[2]);
    float const d4 = dot3(ac[0], ac[1], ac[2], bp[0], bp[1], bp[2]);
    if (d3 >= 0.0f && d4 <= d3) {
        v = 1.0f;
        return;
    }
    float const vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        v = d1 / (d1 - d3);
        return;
    }
    float const cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    float const d5 = dot3(ab[0], ab[1], ab[2], cp[0], cp[1], cp[2]);
    float const d6 = dot3(ac[0], ac[1], ac[2], cp[0], cp[1], cp[2]);
    if (d6 >= 0.0f && d5 <= d6) {
        w = 1.0f;
        return;
    }
    float const vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        w = d2 / (d2 - d6);
        return;
    }
    float const va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        v = 1.0f - w;
        return;
    }
    float const denom = 1.0f / (va + vb + vc);
    v = vb * denom;
    w = vc * denom;
}

// Per-point part of the distance adjoint once the nearest record is known.
__device__ inline void distance_adjoint_finish(
    distance_scene const &sc, float const (&p)[3], std::size_t q, int st, std::uint32_t rec, float uq,
    std::uint32_t const *__restrict__ prim, std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, float *dvx, float *dvy, float *dvz, float *__restrict__ dpx,
    float *__restrict__ dpy, float *__restrict__ dpz
) {
    float g[3] = {0.0f, 0.0f, 0.0f};
    if (st == k_dist_overflow) {
        g[0] = g[1] = g[2] = __builtin_nanf("");
    } else if (st == k_dist_done && rec != 0xffffffffu) {
        tri_record const t = sc.tris[rec];
        float const a[3] = {t.v0x, t.v0y, t.v0z};
        float const b[3] = {t.v0x + (t.v1x - t.v0x), t.v0y + (t.v1y - t.v0y), t.v0z + (t.v1z - t.v0z)};
        float const c[3] = {t.v0x + (t.v2x - t.v0x), t.v0y + (t.v2y - t.v0y), t.v0z + (t.v2z - t.v0z)};
        float v, w;
        closest_barycentrics(p, a, b, c, v, w);
        float const bary[3] = {1.0f - v - w, v, w};
        float r[3];
        for (int k = 0; k < 3; ++k)
            r[k] = p[k] - (bary[0] * a[k] + bary[1] * b[k] + bary[2] * c[k]);
        float const len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (len > 0.0f) {
            float const su = uq / len;
            std::uint32_t const pr = prim[rec];
            std::uint32_t const ids[3] = {i0[pr], i1[pr], i2[pr]};
            for (int k = 0; k < 3; ++k)
                g[k] = su * r[k];
            for (int vtx = 0; vtx < 3; ++vtx) {
                atomicAdd(dvx + ids[vtx], -bary[vtx] * g[0]);
                atomicAdd(dvy + ids[vtx], -bary[vtx] * g[1]);
                atomicAdd(dvz + ids[vtx], -bary[vtx] * g[2]);
            }
        }
    }
    if (dpx != nullptr) {
        dpx[q] = g[0];
        dpy[q] = g[1];
        dpz[q] = g[2];
    }
}

// Adjoint of sum_p u_p d_p: the forward traversal (same decisions) remembers the nearest record;
// with c = sum b_i v_i its closest point and n = (p - c) / |p - c|, d d/dp = n and
// d d/dv_i = -b_i n (the closest point moves with the triangle).  Points on the surface, beyond the
// band or overflowing the stack contribute nothing (NaN dp on overflow).
template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_adjoint_kernel(
    distance_scene sc, float const *__restrict__ px, float const *__restrict__ py, float const *__restrict__ pz,
    float const *__restrict__ u, std::uint32_t const *__restrict__ order, std::size_t n, float culling_band,
    std::uint32_t const *__restrict__ prim, std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, float *dvx, float *dvy, float *dvz, float *__restrict__ dpx,
    float *__restrict__ dpy, float *__restrict__ dpz
) {
    // Four lanes per point (distance_quad_traverse: the nearest record of distance_traverse); the
    // quad's first lane finishes the point.  The grid covers 4 n lanes.
    constexpr int k_quads = k_block / 4;
    __shared__ std::uint32_t s_stack[StackCap * k_quads];
    std::size_t const i = (static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x) >> 2;
    if (i >= n)
        return;
    int const j = static_cast<int>(threadIdx.x & 3);
    std::size_t const q = order ? order[i] : i;
    float const p[3] = {px[q], py[q], pz[q]};
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    std::uint32_t rec = 0xffffffffu;
    int const st = distance_quad_traverse<StackCap>(sc, j, p[0], p[1], p[2], best, rec,
                                                    s_stack + (threadIdx.x >> 2), k_quads);
    if (j == 0)
        distance_adjoint_finish(sc, p, q, st, rec, u[q], prim, i0, i1, i2, dvx, dvy, dvz, dpx, dpy, dpz);
}


// Distance adjoint from the forward's nearest records: no traversal.
__global__ __launch_bounds__(k_block) void distance_adjoint_nearest_kernel(
    distance_scene sc, float const *__restrict__ px, float const *__restrict__ py, float const *__restrict__ pz,
    float const *__restrict__ u, std::uint32_t const *__restrict__ nearest, std::size_t n,
    std::uint32_t const *__restrict__ prim, std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, float *dvx, float *dvy, float *dvz, float *__restrict__ dpx,
    float *__restrict__ dpy, float *__restrict__ dpz
) {
    std::size_t const q = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (q >= n)
        return;
    std::uint32_t const code = nearest[q];
    bool const valid = code < sc.nt;
    int const st = code == k_nearest_overflow ? k_dist_overflow : (valid ? k_dist_done : k_dist_zero);
    float const p[3] = {px[q], py[q], pz[q]};
    distance_adjoint_finish(sc, p, q, st, valid ? code : k_nearest_none, u[q], prim, i0, i1, i2, dvx, dvy, dvz,
                            dpx, dpy, dpz);
}

template <int Cap>
void launch_distance(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, std::uint32_t const *order,
    float *out, std::size_t n, float band, bool wave, bool quad, std::uint32_t *nearest, hipStream_t stream
) {
    distance_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count, bvh.root.reference};
    if (wave && order != nullptr)
        distance_wave_kernel<Cap><<<grid_for(n, k_block), k_block, 0, stream>>>(sc, qx, qy, qz, order, out, n, band,
                                                                                  nearest);
    else if (quad)
        distance_quad_kernel<Cap><<<grid_for(4 * n, k_block), k_block, 0, stream>>>(sc, qx, qy, qz, order, out, n,
                                                                                      band, nearest);
    else
        distance_kernel<Cap><<<grid_for(n, k_block), k_block, 0, stream>>>(sc, qx, qy, qz, order, out, n, band,
                                                                             nearest);
    GWN_LAUNCH_CHECK();
}

} // namespace

void unsigned_distance(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, float *out, std::size_t n,
    float culling_band, query_launch const &cfg, std::uint32_t *nearest, hipStream_t stream
) {
    if (n == 0)
        return;
    bool const sort = cfg.sort_queries > 0 || (cfg.sort_queries < 0 && n >= (1u << 16));
    bool const wave = cfg.sort_queries == 2;
    // Four lanes per point unless one is asked for: the results are the same, and on MI355X the
    // quad traversal measured 2.2-2.8x faster from 65,536 to 16,777,216 points.
    bool const quad = cfg.lanes_per_point != 1;
    device_buffer<std::uint32_t> order;
    if (sort) {
        order.allocate(n, stream);
        query_order(bvh, qx, qy, qz, order.ptr, n, stream);
    }
    int const need = static_cast<int>(bvh.internal_stack_bound);
    if (need <= 32)
        launch_distance<32>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, wave, quad, nearest, stream);
    else if (need <= 64)
        launch_distance<64>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, wave, quad, nearest, stream);
    else if (need <= 128)
        launch_distance<128>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, wave, quad, nearest, stream);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
}

void unsigned_distance_work(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, std::size_t n, float culling_band,
    std::uint32_t *counters, hipStream_t stream
) {
    if (n == 0)
        return;
    distance_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count, bvh.root.reference};
    int const need = static_cast<int>(bvh.internal_stack_bound);
    unsigned const grid = grid_for(n, k_block);
    if (need <= 32)
        distance_work_kernel<32><<<grid, k_block, 0, stream>>>(sc, qx, qy, qz, n, culling_band, counters);
    else if (need <= 64)
        distance_work_kernel<64><<<grid, k_block, 0, stream>>>(sc, qx, qy, qz, n, culling_band, counters);
    else if (need <= 128)
        distance_work_kernel<128><<<grid, k_block, 0, stream>>>(sc, qx, qy, qz, n, culling_band, counters);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
    GWN_LAUNCH_CHECK();
}

void unsigned_distance_adjoint(
    gwn_geometry_handle const &g, gwn_bvh_handle const &bvh, float const *px, float const *py, float const *pz,
    float const *u, std::size_t n, float culling_band, std::uint32_t const *nearest, float *dvx, float *dvy,
    float *dvz, float *dpx, float *dpy, float *dpz, hipStream_t stream
) {
    if (n == 0 || bvh.triangle_count == 0)
        return;
    if (nearest != nullptr) {
        distance_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count,
                                bvh.root.reference};
        distance_adjoint_nearest_kernel<<<grid_for(n, k_block), k_block, 0, stream>>>(
            sc, px, py, pz, u, nearest, n, bvh.primitive_indices.ptr, g.i0.ptr, g.i1.ptr, g.i2.ptr, dvx, dvy, dvz,
            dpx, dpy, dpz);
        GWN_LAUNCH_CHECK();
        return;
    }
    device_buffer<std::uint32_t> order;
    if (n >= (1u << 16)) {
        order.allocate(n, stream);
        query_order(bvh, px, py, pz, order.ptr, n, stream);
    }
    distance_scene const sc{bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count, bvh.root.reference};
    int const need = static_cast<int>(bvh.internal_stack_bound);
#define GWN_DADJ(C)                                                                                            \
    distance_adjoint_kernel<C><<<grid_for(4 * n, k_block), k_block, 0, stream>>>(                             \
        sc, px, py, pz, u, order.ptr, n, culling_band, bvh.primitive_indices.ptr, g.i0.ptr, g.i1.ptr, g.i2.ptr, \
        dvx, dvy, dvz, dpx, dpy, dpz)
    if (need <= 32)
        GWN_DADJ(32);
    else if (need <= 64)
        GWN_DADJ(64);
    else if (need <= 128)
        GWN_DADJ(128);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
#undef GWN_DADJ
    GWN_LAUNCH_CHECK();
}

} // namespace gwn
