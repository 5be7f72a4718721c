// gwn/detail/gwn_device_ray.hpp - per-ray first hit on gfx950: closed-interval slab tests, the
// oriented Pluecker triangle test with the stable normal, and the packed near-first and canonical
// traversals of one ray.
//
// Reference: include/gwn/detail/gwn_query_ray_impl.cuh:20-595, include/gwn/gwn_query.cuh:63-95.
//
// Shared by libgwn_hip.so's ray kernels and the public device point API (gwn/gwn_device.hpp).
// Arithmetic is unfused (every function body sets `#pragma clang fp contract(off)` for itself), so
// records equal the batch's and the CPU oracle's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "gwn_device_layout.hpp"

namespace gwn::detail::ray {

constexpr std::uint32_t k_payload_leaf = 0x80000000u;

// gwn_ray_first_hit_result<float, uint32_t> (gwn_query.cuh:74-95), 32 bytes.
struct ray_hit {
    float t, u, v, nx, ny, nz;
    std::uint32_t primitive_id;
    std::uint8_t status; // 0 miss, 1 hit, 2 overflow
    std::uint8_t pad[3];
};
static_assert(sizeof(ray_hit) == 32, "ray hit record must be 32 bytes");

__device__ inline float min_std(float a, float b) { return (b < a) ? b : a; } // std::min
__device__ inline float max_std(float a, float b) { return (a < b) ? b : a; } // std::max

struct ray_state {
    float o[3], d[3], inv[3];
    bool zero[3];
    bool parallel;
};

// One closed slab; parallel axes test containment and leave the interval alone (:82-98).
__device__ inline bool axis_update(float o, float lo, float hi, float inv, bool zero, float &t_near, float &t_far) {
#pragma clang fp contract(off)
    if (zero)
        return o >= lo && o <= hi;
    float const t0 = (lo - o) * inv;
    float const t1 = (hi - o) * inv;
    t_near = max_std(t_near, min_std(t0, t1));
    t_far = min_std(t_far, max_std(t0, t1));
    return t_near <= t_far;
}

// gwn_ray_aabb_intersect_interval_impl (:100-140).
__device__ inline bool box_canonical(ray_state const &r, aabb const &b, float t_min, float t_max, float &t_near_out) {
#pragma clang fp contract(off)
    if (!(t_max >= t_min))
        return false;
    float t_near = t_min, t_far = t_max;
    if (!axis_update(r.o[0], b.min_x, b.max_x, r.inv[0], r.zero[0], t_near, t_far) ||
        !axis_update(r.o[1], b.min_y, b.max_y, r.inv[1], r.zero[1], t_near, t_far) ||
        !axis_update(r.o[2], b.min_z, b.max_z, r.inv[2], r.zero[2], t_near, t_far))
        return false;
    if (t_far < t_near)
        return false;
    t_near_out = t_near;
    return true;
}

// Packed traversal's box test (:257-290): branch-free slabs unless an axis is parallel.
__device__ inline bool box_packed(ray_state const &r, aabb const &b, float t_min, float t_best, float &t_near_out) {
#pragma clang fp contract(off)
    if (!r.parallel) {
        float const tx0 = (b.min_x - r.o[0]) * r.inv[0], tx1 = (b.max_x - r.o[0]) * r.inv[0];
        float const ty0 = (b.min_y - r.o[1]) * r.inv[1], ty1 = (b.max_y - r.o[1]) * r.inv[1];
        float const tz0 = (b.min_z - r.o[2]) * r.inv[2], tz1 = (b.max_z - r.o[2]) * r.inv[2];
        float const t_near =
            max_std(t_min, max_std(min_std(tx0, tx1), max_std(min_std(ty0, ty1), min_std(tz0, tz1))));
        float const t_far =
            min_std(t_best, min_std(max_std(tx0, tx1), min_std(max_std(ty0, ty1), max_std(tz0, tz1))));
        t_near_out = t_near;
        return t_near <= t_far;
    }
    float t_near = t_min, t_far = t_best;
    if (!axis_update(r.o[0], b.min_x, b.max_x, r.inv[0], r.zero[0], t_near, t_far) ||
        !axis_update(r.o[1], b.min_y, b.max_y, r.inv[1], r.zero[1], t_near, t_far) ||
        !axis_update(r.o[2], b.min_z, b.max_z, r.inv[2], r.zero[2], t_near, t_far))
        return false;
    t_near_out = t_near;
    return true;
}

__device__ inline void cross3(float const *a, float const *b, float *c) {
#pragma clang fp contract(off)
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline float dot3(float const *a, float const *b) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Component-wise stable oriented normal (gwn_stable_triangle_normal_impl, :143-172).
__device__ inline void stable_normal(float const *a, float const *b, float const *c, float *n) {
#pragma clang fp contract(off)
    float const abx = a[2] * b[1], aby = a[0] * b[2], abz = a[1] * b[0];
    float const bcx = b[2] * c[1], bcy = b[0] * c[2], bcz = b[1] * c[0];
    n[0] = (fabsf(abx) < fabsf(bcx)) ? a[1] * b[2] - abx : b[1] * c[2] - bcx;
    n[1] = (fabsf(aby) < fabsf(bcy)) ? a[2] * b[0] - aby : b[2] * c[0] - bcy;
    n[2] = (fabsf(abz) < fabsf(bcz)) ? a[0] * b[1] - abz : b[0] * c[1] - bcz;
}

struct best_hit {
    float t, u, v, n[3];
    std::uint32_t prim;
    bool found;
};

// Oriented Pluecker edge functions (gwn_ray_triangle_intersect_impl, :175-239), on the record
// rebuilt in the reference's (v0, v0 + e1, v0 + e2) form.
__device__ inline void intersect_triangle(
    ray_state const &r, tri_record const &tr, std::uint32_t prim, float t_min, best_hit &best
) {
#pragma clang fp contract(off)
    float const e1x = tr.v1x - tr.v0x, e1y = tr.v1y - tr.v0y, e1z = tr.v1z - tr.v0z;
    float const e2x = tr.v2x - tr.v0x, e2y = tr.v2y - tr.v0y, e2z = tr.v2z - tr.v0z;
    float const v1[3] = {tr.v0x + e1x, tr.v0y + e1y, tr.v0z + e1z};
    float const v2[3] = {tr.v0x + e2x, tr.v0y + e2y, tr.v0z + e2z};
    float const p0[3] = {tr.v0x - r.o[0], tr.v0y - r.o[1], tr.v0z - r.o[2]};
    float const p1[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    float const p2[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    float const e0[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    float const ea[3] = {p0[0] - p1[0], p0[1] - p1[1], p0[2] - p1[2]};
    float const eb[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
    float s[3], c[3];
    s[0] = p2[0] + p0[0]; s[1] = p2[1] + p0[1]; s[2] = p2[2] + p0[2];
    cross3(e0, s, c);
    float const u = dot3(c, r.d);
    s[0] = p0[0] + p1[0]; s[1] = p0[1] + p1[1]; s[2] = p0[2] + p1[2];
    cross3(ea, s, c);
    float const v = dot3(c, r.d);
    s[0] = p1[0] + p2[0]; s[1] = p1[1] + p2[1]; s[2] = p1[2] + p2[2];
    cross3(eb, s, c);
    float const w = dot3(c, r.d);
    float const uvw = u + v + w;
    float const eps = FLT_EPSILON * fabsf(uvw);
    float const mn = min_std(u, min_std(v, w));
    float const mx = max_std(u, max_std(v, w));
    if (!(mn >= -eps || mx <= eps))
        return;
    float ng[3];
    stable_normal(e0, ea, eb, ng);
    float const den = 2.0f * dot3(ng, r.d);
    if (den == 0.0f)
        return;
    float const t = (2.0f * dot3(p0, ng)) / den;
    if (!(t_min <= t && t <= best.t))
        return;
    float bu, bv;
    if (fabsf(uvw) >= FLT_MIN) {
        float const inv = 1.0f / uvw;
        bu = u * inv;
        bv = v * inv;
    } else {
        if (uvw == 0.0f)
            return;
        bu = u / uvw;
        bv = v / uvw;
    }
    if (best.found && !(t < best.t))
        return;
    best.t = t;
    best.u = bu;
    best.v = bv;
    best.n[0] = ng[0];
    best.n[1] = ng[1];
    best.n[2] = ng[2];
    best.prim = prim;
    best.found = true;
}

struct ray_scene {
    bvh_node4 const *nodes;
    std::size_t node_count;
    tri_record const *tris;
    std::uint32_t const *prim;
    std::size_t nt;
};

__device__ inline void visit_leaf(ray_scene const &sc, ray_state const &r, std::uint64_t ref, float t_min, best_hit &best,
                                  traversal_work *work = nullptr) {
#pragma clang fp contract(off)
    std::uint64_t const off = ref_offset(ref);
    std::uint32_t const cnt = ref_count(ref);
    for (std::uint32_t k = 0; k < cnt; ++k) {
        if (off + k >= sc.nt)
            continue;
        if (work != nullptr)
            ++work->tris;
        intersect_triangle(r, sc.tris[off + k], sc.prim[off + k], t_min, best);
    }
}


// Packed near-first traversal (:242-358).  Returns true on stack overflow.
template <int Cap>
__device__ inline bool traverse_packed(ray_scene const &sc, std::uint32_t root, ray_state const &r, float t_min, best_hit &best,
                                       traversal_work *work = nullptr) {
#pragma clang fp contract(off)
    // One loop exit, at the head: no return from inside the unrolled push loop (the gfx950 backend
    // of ROCm 7.2 has lowered such early exits wrongly).
    std::uint32_t stack[Cap];
    int sp = 0;
    bool overflow = false, more = true;
    std::uint32_t payload = root;
    while (more && !overflow) {
        bool descended = false;
        if (payload & k_payload_leaf) {
            int const slot = static_cast<int>(payload & 3u);
            std::uint32_t const parent = (payload & ~k_payload_leaf) >> 2;
            visit_leaf(sc, r, sc.nodes[parent].children[slot].reference, t_min, best, work);
        } else if (payload < sc.node_count) {
            bvh_node4 const &node = sc.nodes[payload];
            float et[4];
            std::uint32_t ep[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                et[s] = INFINITY;
                ep[s] = 0u;
                bvh_child const ch = node.children[s];
                if (!ref_valid(ch.reference))
                    continue;
                bool const leaf = ref_leaf(ch.reference);
                if (!leaf && ref_offset(ch.reference) >= sc.node_count)
                    continue;
                if (work != nullptr)
                    ++work->boxes;
                float t_near;
                if (!box_packed(r, ch.bounds, t_min, best.t, t_near))
                    continue;
                et[s] = t_near;
                ep[s] = leaf ? (k_payload_leaf | (payload << 2) | static_cast<std::uint32_t>(s))
                             : static_cast<std::uint32_t>(ref_offset(ch.reference));
            }
            sort4(et, ep);
#pragma unroll
            for (int s = 3; s > 0; --s) {
                if (et[s] == INFINITY || overflow)
                    continue;
                if (sp < Cap)
                    stack[sp++] = ep[s];
                else
                    overflow = true;
            }
            if (et[0] != INFINITY) {
                payload = ep[0];
                descended = true;
            }
        }
        if (!descended) {
            if (sp > 0)
                payload = stack[--sp];
            else
                more = false;
        }
    }
    return overflow;
}

// Canonical traversal (:485-560).  Returns true on stack overflow.
template <int Cap>
__device__ inline bool traverse_canonical(ray_scene const &sc, std::uint64_t root_ref, ray_state const &r, float t_min,
                                          best_hit &best, traversal_work *work = nullptr) {
#pragma clang fp contract(off)
    std::uint32_t stack[Cap];
    int sp = 0;
    bool overflow = false;
    std::uint64_t reference = root_ref;
    while (reference != 0 && !overflow) {
        std::uint64_t next = 0;
        if (ref_leaf(reference)) {
            visit_leaf(sc, r, reference, t_min, best, work);
        } else if (ref_internal(reference) && ref_offset(reference) < sc.node_count) {
            bvh_node4 const &node = sc.nodes[ref_offset(reference)];
            float et[4];
            std::uint64_t er[4];
            int m = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bvh_child const ch = node.children[s];
                if (!ref_valid(ch.reference))
                    continue;
                if (work != nullptr)
                    ++work->boxes;
                float t_near;
                if (!box_canonical(r, ch.bounds, t_min, best.t, t_near))
                    continue;
                et[m] = t_near;
                er[m] = ch.reference;
                ++m;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s >= m) {
                    et[s] = INFINITY;
                    er[s] = 0;
                }
            }
            sort4(et, er);
#pragma unroll
            for (int s = 0; s < 4; ++s)
                if (s < m && ref_leaf(er[s]) && et[s] <= best.t)
                    visit_leaf(sc, r, er[s], t_min, best, work);
#pragma unroll
            for (int s = 3; s >= 0; --s) {
                if (s >= m || !ref_internal(er[s]) || et[s] > best.t || overflow)
                    continue;
                if (next != 0) {
                    if (sp < Cap)
                        stack[sp++] = static_cast<std::uint32_t>(ref_offset(next));
                    else
                        overflow = true;
                }
                next = er[s];
            }
        }
        if (next == 0 && sp > 0)
            next = encode_internal(stack[--sp]);
        reference = next;
    }
    return overflow;
}


// One ray's first hit (the batch kernel's per-ray body): direction precompute, the packed
// near-first or the canonical traversal, and the 32-byte record (t = -1 and primitive 0xFFFFFFFF
// on a miss, status 2 on stack overflow with the best hit found so far).  A zero direction misses.
// With `work`, the traversal's slab and triangle tests are tallied there (zero for a zero
// direction).
template <bool Packed, int Cap>
__device__ inline ray_hit ray_first_hit_lane(
    ray_scene const &sc, std::uint64_t root_ref, float ox, float oy, float oz, float dx, float dy, float dz,
    float t_min, float t_max, traversal_work *work
) {
#pragma clang fp contract(off)
    ray_hit res{};
    res.t = -1.0f;
    res.primitive_id = 0xffffffffu;
    ray_state r;
    r.o[0] = ox;
    r.o[1] = oy;
    r.o[2] = oz;
    r.d[0] = dx;
    r.d[1] = dy;
    r.d[2] = dz;
    if (work != nullptr)
        *work = traversal_work{};
    float const len2 = r.d[0] * r.d[0] + r.d[1] * r.d[1] + r.d[2] * r.d[2];
    if (len2 > 0.0f) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            r.zero[k] = r.d[k] == 0.0f;
            r.inv[k] = r.zero[k] ? 0.0f : 1.0f / r.d[k];
        }
        r.parallel = r.zero[0] || r.zero[1] || r.zero[2];
        best_hit best{t_max, 0.0f, 0.0f, {0.0f, 0.0f, 0.0f}, 0xffffffffu, false};
        bool overflow;
        if constexpr (Packed)
            overflow = traverse_packed<Cap>(sc, static_cast<std::uint32_t>(ref_offset(root_ref)), r, t_min, best, work);
        else
            overflow = traverse_canonical<Cap>(sc, root_ref, r, t_min, best, work);
        if (best.found) {
            res.t = best.t;
            res.u = best.u;
            res.v = best.v;
            res.nx = best.n[0];
            res.ny = best.n[1];
            res.nz = best.n[2];
            res.primitive_id = best.prim;
            res.status = 1;
        }
        if (overflow)
            res.status = 2;
    }
    return res;
}

} // namespace gwn::detail::ray
