// gwn/detail/gwn_device_distance.hpp - per-point unsigned distance on gfx950: point-triangle and
// point-box squared distances and the near-first culled traversal of one point.
//
// Reference: include/gwn/detail/gwn_query_distance_impl.cuh:20-192,
// include/gwn/detail/gwn_query_geometry_impl.cuh:46-114.
//
// Shared by libgwn_hip.so's distance kernels and the public device point API (gwn/gwn_device.hpp).
// Arithmetic is unfused (every function body sets `#pragma clang fp contract(off)` for itself) and
// the square root correctly rounded, so results equal the batch's and the CPU oracle's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "gwn_device_layout.hpp"

namespace gwn::detail::distance {

__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
#pragma clang fp contract(off)
    return ax * bx + ay * by + az * bz;
}

// Squared distance from p to the closed triangle (a, b, c) by Voronoi regions (Ericson).
__device__ inline float point_triangle_distance2(
    float px, float py, float pz, float ax, float ay, float az, float bx, float by, float bz, float cx,
    float cy, float cz
) {
#pragma clang fp contract(off)
    float const abx = bx - ax, aby = by - ay, abz = bz - az;
    float const acx = cx - ax, acy = cy - ay, acz = cz - az;
    float const apx = px - ax, apy = py - ay, apz = pz - az;
    float const d1 = dot3(abx, aby, abz, apx, apy, apz);
    float const d2 = dot3(acx, acy, acz, apx, apy, apz);
    if (d1 <= 0.0f && d2 <= 0.0f)
        return dot3(apx, apy, apz, apx, apy, apz);
    float const bpx = px - bx, bpy = py - by, bpz = pz - bz;
    float const d3 = dot3(abx, aby, abz, bpx, bpy, bpz);
    float const d4 = dot3(acx, acy, acz, bpx, bpy, bpz);
    if (d3 >= 0.0f && d4 <= d3)
        return dot3(bpx, bpy, bpz, bpx, bpy, bpz);
    float const vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float const v = d1 / (d1 - d3);
        float const rx = px - (ax + v * abx), ry = py - (ay + v * aby), rz = pz - (az + v * abz);
        return dot3(rx, ry, rz, rx, ry, rz);
    }
    float const cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    float const d5 = dot3(abx, aby, abz, cpx, cpy, cpz);
    float const d6 = dot3(acx, acy, acz, cpx, cpy, cpz);
    if (d6 >= 0.0f && d5 <= d6)
        return dot3(cpx, cpy, cpz, cpx, cpy, cpz);
    float const vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float const w = d2 / (d2 - d6);
        float const rx = px - (ax + w * acx), ry = py - (ay + w * acy), rz = pz - (az + w * acz);
        return dot3(rx, ry, rz, rx, ry, rz);
    }
    float const va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float const w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        float const rx = px - (bx + w * (cx - bx)), ry = py - (by + w * (cy - by)), rz = pz - (bz + w * (cz - bz));
        return dot3(rx, ry, rz, rx, ry, rz);
    }
    float const denom = 1.0f / (va + vb + vc);
    float const v = vb * denom, w = vc * denom;
    float const rx = px - (ax + v * abx + w * acx);
    float const ry = py - (ay + v * aby + w * acy);
    float const rz = pz - (az + v * abz + w * acz);
    return dot3(rx, ry, rz, rx, ry, rz);
}

__device__ inline float max0(float x) { return (x < 0.0f) ? 0.0f : x; } // std::max(x, 0)

__device__ inline float box_distance2(float qx, float qy, float qz, aabb const &b) {
#pragma clang fp contract(off)
    float const dx = max0(b.min_x - qx) + max0(qx - b.max_x);
    float const dy = max0(b.min_y - qy) + max0(qy - b.max_y);
    float const dz = max0(b.min_z - qz) + max0(qz - b.max_z);
    return dx * dx + dy * dy + dz * dz;
}

// Leaf range: records rebuilt in the reference's (v0, v0 + e1, v0 + e2) form.  `best_rec`, when
// non-null, receives the record of each strict improvement.
__device__ inline bool visit_leaf(
    tri_record const *__restrict__ tris, std::size_t nt, std::uint64_t ref, float qx, float qy, float qz,
    float &best, std::uint32_t *best_rec, traversal_work *work = nullptr
) {
#pragma clang fp contract(off)
    std::uint64_t const off = ref_offset(ref);
    std::uint32_t const cnt = ref_count(ref);
    for (std::uint32_t k = 0; k < cnt; ++k) {
        if (off + k >= nt)
            continue;
        if (work != nullptr)
            ++work->tris;
        tri_record const t = tris[off + k];
        float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
        float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
        float const d2 = point_triangle_distance2(qx, qy, qz, t.v0x, t.v0y, t.v0z, t.v0x + e1x, t.v0y + e1y,
                                                  t.v0z + e1z, t.v0x + e2x, t.v0y + e2y, t.v0z + e2z);
        if (d2 < best) {
            best = d2;
            if (best_rec != nullptr)
                *best_rec = static_cast<std::uint32_t>(off + k);
        }
    }
    return best > 0.0f;
}

struct distance_scene {
    bvh_node4 const *nodes;
    std::size_t node_count;
    tri_record const *tris;
    std::size_t nt;
    std::uint64_t root_ref;
};

enum : int { k_dist_done = 0, k_dist_zero = 1, k_dist_overflow = 2 };

// Nearest-record code of a finished traversal (the adjoint's input).
__device__ inline std::uint32_t nearest_code(int st, std::uint32_t rec) {
#pragma clang fp contract(off)
    return st == k_dist_overflow ? k_nearest_overflow : (st == k_dist_done ? rec : k_nearest_none);
}

// Near-first culled traversal; `best` starts at the squared band and ends at the squared
// distance.  Returns k_dist_zero when a triangle touches the query, k_dist_overflow when the
// stack would overflow.
template <int StackCap>
__device__ inline int distance_traverse(
    distance_scene const &sc, float x, float y, float z, float &best, std::uint32_t *best_rec,
    traversal_work *work = nullptr
) {
#pragma clang fp contract(off)
    // One loop exit, at the head (status flags instead of returns inside the unrolled loops): the
    // gfx950 backend (ROCm 7.2) has lowered an early exit inside such a loop wrongly.
    std::uint32_t stack[StackCap];
    int sp = 0;
    int status = k_dist_done;
    std::uint64_t reference = sc.root_ref;
    while (reference != 0 && status == k_dist_done) {
        std::uint64_t next = 0;
        if (ref_leaf(reference)) {
            if (!visit_leaf(sc.tris, sc.nt, reference, x, y, z, best, best_rec, work))
                status = k_dist_zero;
        } else if (ref_internal(reference) && ref_offset(reference) < sc.node_count) {
            bvh_node4 const &node = sc.nodes[ref_offset(reference)];
            float cd[4];
            std::uint64_t cr[4];
            int m = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bvh_child const ch = node.children[s];
                if (!ref_valid(ch.reference))
                    continue;
                if (work != nullptr)
                    ++work->boxes;
                float const d2 = box_distance2(x, y, z, ch.bounds);
                if (d2 >= best)
                    continue;
                cd[m] = d2;
                cr[m] = ch.reference;
                ++m;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s >= m) {
                    cd[s] = INFINITY;
                    cr[s] = 0;
                }
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
            };
            cswap(0, 1);
            cswap(2, 3);
            cswap(0, 2);
            cswap(1, 3);
            cswap(1, 2);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (status == k_dist_done && s < m && ref_leaf(cr[s]) && cd[s] < best &&
                    !visit_leaf(sc.tris, sc.nt, cr[s], x, y, z, best, best_rec, work))
                    status = k_dist_zero;
            }
#pragma unroll
            for (int s = 3; s >= 0; --s) {
                if (status != k_dist_done || s >= m || !ref_internal(cr[s]) || cd[s] >= best)
                    continue;
                if (next != 0) {
                    if (sp < StackCap)
                        stack[sp++] = static_cast<std::uint32_t>(ref_offset(next));
                    else
                        status = k_dist_overflow;
                }
                next = cr[s];
            }
        }
        if (next == 0 && sp > 0)
            next = encode_internal(stack[--sp]);
        reference = next;
    }
    return status;
}

} // namespace gwn::detail::distance
