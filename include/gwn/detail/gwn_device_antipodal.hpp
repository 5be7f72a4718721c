// gwn/detail/gwn_device_antipodal.hpp - per-query antipodal winding number and gradient on gfx950:
// projected-edge classification with the symbolic-perturbation sign, signed crossings of the
// positive coordinate ray by BVH traversal, and the boundary chain's signed spherical areas
// (Kahan sums in FP32, or FP64 areas with the precision split).
//
// Reference: include/gwn/detail/gwn_query_winding_antipodal_impl.cuh:55-552.
//
// Shared by libgwn_hip.so's antipodal kernels and the public device point API
// (gwn/gwn_device.hpp).  Arithmetic is unfused: every function body that multiplies sets
// `#pragma clang fp contract(off)` for itself, so results equal the batch's whatever the
// includer's contraction mode.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../gwn_hip.h"
#include "gwn_device_layout.hpp"

namespace gwn::detail::antipodal {

constexpr int k_axis_x = 0, k_axis_y = 1, k_axis_z = 2;
constexpr float k_pi_f = 3.14159265358979323846f;
constexpr double k_pi_d = 3.14159265358979323846;

struct v3 {
    float x, y, z;
};

__device__ inline float comp(v3 const &v, int axis) { return axis == k_axis_x ? v.x : (axis == k_axis_y ? v.y : v.z); }
__device__ inline float dot(v3 const &a, v3 const &b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ inline v3 cross(v3 const &a, v3 const &b) {
#pragma clang fp contract(off)
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline v3 sub(v3 const &a, v3 const &b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }

__device__ inline int sign_of(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }

// Projected edge class (:64-110): perturbed sign; singular when the segment holds the query.
__device__ inline int classify_edge(v3 const &s, v3 const &e, int axis, bool &singular, float &det) {
#pragma clang fp contract(off)
    float ax, ay, bx, by;
    if (axis == k_axis_x) {
        ax = s.y; ay = s.z; bx = e.y; by = e.z;
    } else if (axis == k_axis_y) {
        ax = s.z; ay = s.x; bx = e.z; by = e.x;
    } else {
        ax = s.x; ay = s.y; bx = e.x; by = e.y;
    }
    float const d = ax * by - ay * bx;
    int sg = sign_of(d);
    if (sg == 0) {
        sg = sign_of(ay - by);
        if (sg == 0)
            sg = sign_of(bx - ax);
    }
    singular = d == 0.0f && (ax * bx + ay * by <= 0.0f);
    det = d;
    return sg;
}

// One triangle's crossing (:122-176): returns true when singular.
__device__ inline bool triangle_crossing(v3 const &o, int axis, tri_record const &t, float &crossing) {
#pragma clang fp contract(off)
    float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    v3 const a{t.v0x, t.v0y, t.v0z};
    v3 const b{t.v0x + e1x, t.v0y + e1y, t.v0z + e1z};
    v3 const c{t.v0x + e2x, t.v0y + e2y, t.v0z + e2z};
    v3 const pa = sub(a, o), pb = sub(b, o), pc = sub(c, o);
    v3 const n = cross(sub(pb, pa), sub(pc, pa));
    float const na = comp(n, axis);
    if (na == 0.0f)
        return false;
    bool s1, s2, s3;
    float d;
    int const c1 = classify_edge(pa, pb, axis, s1, d);
    int const c2 = classify_edge(pb, pc, axis, s2, d);
    int const c3 = classify_edge(pc, pa, axis, s3, d);
    if (s1 || s2 || s3 || c1 == 0 || c2 == 0 || c3 == 0)
        return true;
    if (!((c1 >= 0 && c2 >= 0 && c3 >= 0) || (c1 <= 0 && c2 <= 0 && c3 <= 0)))
        return false;
    float const tt = dot(n, pa) / na;
    if (tt == 0.0f)
        return true;
    if (!(tt > 0.0f))
        return false;
    crossing = na > 0.0f ? 1.0f : -1.0f;
    return false;
}

__device__ inline bool axis_ray_enters(v3 const &o, int axis, aabb const &b) {
#pragma clang fp contract(off)
    if (axis == k_axis_x)
        return o.y >= b.min_y && o.y <= b.max_y && o.z >= b.min_z && o.z <= b.max_z && b.max_x >= o.x;
    if (axis == k_axis_y)
        return o.z >= b.min_z && o.z <= b.max_z && o.x >= b.min_x && o.x <= b.max_x && b.max_y >= o.y;
    return o.x >= b.min_x && o.x <= b.max_x && o.y >= b.min_y && o.y <= b.max_y && b.max_z >= o.z;
}

struct crossing_scene {
    bvh_node4 const *nodes;
    std::size_t node_count;
    tri_record const *tris;
    std::size_t nt;
    std::uint64_t root;
};

__device__ inline bool leaf_crossings(crossing_scene const &sc, std::uint64_t ref, v3 const &o, int axis, float &sum) {
#pragma clang fp contract(off)
    std::uint64_t const off = ref_offset(ref);
    std::uint32_t const cnt = ref_count(ref);
    bool singular = false;
    for (std::uint32_t k = 0; k < cnt && !singular; ++k) {
        if (off + k >= sc.nt)
            continue;
        float cr = 0.0f;
        singular = triangle_crossing(o, axis, sc.tris[off + k], cr);
        sum += singular ? 0.0f : cr;
    }
    return singular;
}

// Signed crossing count (:392-490): returns true when singular; NaN sum on stack overflow.
//
// The loop has one exit, at its head: singular triangles and stack overflow set flags instead of
// returning from inside the unrolled child loop.  With early returns there, the gfx950 backend
// (ROCm 7.2 LLVM) copied the overflow path's `next` into every lane that had just pushed, so
// those lanes revisited the pushed node and skipped the new child; the LLVM IR was correct.
template <int Cap>
__device__ inline bool crossing_count(crossing_scene const &sc, v3 const &o, int axis, float &sum) {
#pragma clang fp contract(off)
    sum = 0.0f;
    if (ref_leaf(sc.root))
        return leaf_crossings(sc, sc.root, o, axis, sum);
    constexpr std::uint32_t k_none = 0xffffffffu;
    std::uint32_t stack[Cap];
    int sp = 0;
    bool singular = false, overflow = false;
    std::uint32_t node = static_cast<std::uint32_t>(ref_offset(sc.root));
    while (node != k_none && !singular && !overflow) {
        std::uint32_t next = k_none;
        if (node < sc.node_count) {
            bvh_node4 const &nd = sc.nodes[node];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bvh_child const ch = nd.children[s];
                bool const enter = ref_valid(ch.reference) && axis_ray_enters(o, axis, ch.bounds);
                if (enter && ref_leaf(ch.reference)) {
                    singular = singular || leaf_crossings(sc, ch.reference, o, axis, sum);
                } else if (enter) {
                    if (next != k_none) {
                        if (sp < Cap)
                            stack[sp++] = next;
                        else
                            overflow = true;
                    }
                    next = static_cast<std::uint32_t>(ref_offset(ch.reference));
                }
            }
        }
        if (next == k_none && sp > 0)
            next = stack[--sp];
        node = next;
    }
    if (overflow) {
        sum = __builtin_nanf("");
        return false;
    }
    return singular;
}

__device__ inline v3 minus_ray(int axis) {
#pragma clang fp contract(off)
    return v3{-(axis == k_axis_x ? 1.0f : 0.0f), -(axis == k_axis_y ? 1.0f : 0.0f), -(axis == k_axis_z ? 1.0f : 0.0f)};
}

// Signed spherical area of (-ray, s, e) (:178-204), reference arithmetic.
__device__ inline float spherical_area_f32(v3 const &m, v3 const &s, v3 const &e, float det) {
#pragma clang fp contract(off)
    float const sn = sqrtf(dot(s, s)), en = sqrtf(dot(e, e));
    if (!(sn > 0.0f) || !(en > 0.0f))
        return __builtin_nanf("");
    float num = -det;
    if (det == 0.0f)
        num = dot(m, cross(s, e));
    float const den = sn * en + dot(m, s) * en + dot(s, e) + dot(e, m) * sn;
    return 2.0f * atan2f(num, den);
}

__device__ inline double spherical_area_f64(double3 const &m, double3 const &s, double3 const &e) {
#pragma clang fp contract(off)
    double const sn = sqrt(s.x * s.x + s.y * s.y + s.z * s.z), en = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    if (!(sn > 0.0) || !(en > 0.0))
        return __builtin_nan("");
    double const cx = s.y * e.z - s.z * e.y, cy = s.z * e.x - s.x * e.z, cz = s.x * e.y - s.y * e.x;
    double const num = m.x * cx + m.y * cy + m.z * cz;
    double const den = sn * en + (m.x * s.x + m.y * s.y + m.z * s.z) * en + (s.x * e.x + s.y * e.y + s.z * e.z) +
                       (e.x * m.x + e.y * m.y + e.z * m.z) * sn;
    return 2.0 * atan2(num, den);
}

// d/dq of one area term (:257-334).
__device__ inline v3 area_gradient_f32(v3 const &m, v3 const &s, v3 const &e, float det) {
#pragma clang fp contract(off)
    float const nan = __builtin_nanf("");
    float const sn = sqrtf(dot(s, s)), en = sqrtf(dot(e, e));
    if (sn == 0.0f || en == 0.0f)
        return v3{nan, nan, nan};
    float const is = 1.0f / sn, ie = 1.0f / en;
    v3 const su{s.x * is, s.y * is, s.z * is};
    v3 const eu{e.x * ie, e.y * ie, e.z * ie};
    float num = -det;
    if (det == 0.0f)
        num = dot(m, cross(s, e));
    float const den = sn * en + dot(m, s) * en + dot(m, e) * sn + dot(s, e);
    float const d2 = num * num + den * den;
    if (d2 == 0.0f)
        return v3{nan, nan, nan};
    float const cn = 2.0f * den / d2;
    float const cd = -2.0f * num / d2;
    v3 const c1 = cross(e, m), c2 = cross(m, s);
    float const f1 = 1.0f + dot(m, eu), f2 = 1.0f + dot(m, su);
    v3 g;
    {
        float const gn = -(c1.x + c2.x);
        float const gd = -(en * (m.x + eu.x + f1 * su.x) + sn * (m.x + su.x + f2 * eu.x));
        g.x = cn * gn + cd * gd;
    }
    {
        float const gn = -(c1.y + c2.y);
        float const gd = -(en * (m.y + eu.y + f1 * su.y) + sn * (m.y + su.y + f2 * eu.y));
        g.y = cn * gn + cd * gd;
    }
    {
        float const gn = -(c1.z + c2.z);
        float const gd = -(en * (m.z + eu.z + f1 * su.z) + sn * (m.z + su.z + f2 * eu.z));
        g.z = cn * gn + cd * gd;
    }
    return g;
}

__device__ inline double3 area_gradient_f64(double3 const &m, double3 const &s, double3 const &e) {
#pragma clang fp contract(off)
    double const nan = __builtin_nan("");
    double const sn = sqrt(s.x * s.x + s.y * s.y + s.z * s.z), en = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
    if (sn == 0.0 || en == 0.0)
        return double3{nan, nan, nan};
    double3 const su{s.x / sn, s.y / sn, s.z / sn};
    double3 const eu{e.x / en, e.y / en, e.z / en};
    double const num = m.x * (s.y * e.z - s.z * e.y) + m.y * (s.z * e.x - s.x * e.z) + m.z * (s.x * e.y - s.y * e.x);
    double const den = sn * en + (m.x * s.x + m.y * s.y + m.z * s.z) * en + (m.x * e.x + m.y * e.y + m.z * e.z) * sn +
                       (s.x * e.x + s.y * e.y + s.z * e.z);
    double const d2 = num * num + den * den;
    if (d2 == 0.0)
        return double3{nan, nan, nan};
    double const cn = 2.0 * den / d2, cd = -2.0 * num / d2;
    double const c1x = e.y * m.z - e.z * m.y, c1y = e.z * m.x - e.x * m.z, c1z = e.x * m.y - e.y * m.x;
    double const c2x = m.y * s.z - m.z * s.y, c2y = m.z * s.x - m.x * s.z, c2z = m.x * s.y - m.y * s.x;
    double const f1 = 1.0 + (m.x * eu.x + m.y * eu.y + m.z * eu.z);
    double const f2 = 1.0 + (m.x * su.x + m.y * su.y + m.z * su.z);
    double3 g;
    g.x = cn * -(c1x + c2x) + cd * -(en * (m.x + eu.x + f1 * su.x) + sn * (m.x + su.x + f2 * eu.x));
    g.y = cn * -(c1y + c2y) + cd * -(en * (m.y + eu.y + f1 * su.y) + sn * (m.y + su.y + f2 * eu.y));
    g.z = cn * -(c1z + c2z) + cd * -(en * (m.z + eu.z + f1 * su.z) + sn * (m.z + su.z + f2 * eu.z));
    return g;
}

// A boundary edge staged in LDS: endpoints and the multiplicity as the working Real.
struct staged_edge {
    float sx, sy, sz, ex, ey, ez, mult;
};

struct boundary_view {
    float const *vx, *vy, *vz;
    std::uint32_t const *start, *end;
    std::uint64_t const *mult;
    std::size_t count;
};

// Per-lane accumulation of one axis pass over the chain.
template <int Prec, bool Grad> struct boundary_acc {
    float sum = 0.0f, comp = 0.0f;            // plain winding (Kahan)
    float gs[3] = {0, 0, 0}, gc[3] = {0, 0, 0}; // plain gradient (vector Kahan)
    double sum_d = 0.0;                       // split winding
    double gd[3] = {0.0, 0.0, 0.0};           // split gradient
    bool singular = false;

    __device__ void add(staged_edge const &ed, v3 const &q, int axis) {
#pragma clang fp contract(off)
        v3 const s{ed.sx - q.x, ed.sy - q.y, ed.sz - q.z};
        v3 const e{ed.ex - q.x, ed.ey - q.y, ed.ez - q.z};
        bool sing;
        float det;
        int const sg = classify_edge(s, e, axis, sing, det);
        if (sing) {
            singular = true;
            return;
        }
        v3 const m = minus_ray(axis);
        if constexpr (Prec == GWN_PRECISION_PLAIN) {
            if constexpr (Grad) {
                v3 const g = area_gradient_f32(m, s, e, det);
                float const t[3] = {ed.mult * g.x, ed.mult * g.y, ed.mult * g.z};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    float const ct = t[k] - gc[k];
                    float const next = gs[k] + ct;
                    gc[k] = (next - gs[k]) - ct;
                    gs[k] = next;
                }
            } else {
                float area = spherical_area_f32(m, s, e, det);
                if (axis != k_axis_z && sg > 0 && area < 0.0f)
                    area += 2.0f * k_pi_f;
                float const term = ed.mult * area;
                float const ct = term - comp;
                float const next = sum + ct;
                comp = (next - sum) - ct;
                sum = next;
            }
        } else {
            double3 const md{m.x, m.y, m.z};
            double3 const sd{double(ed.sx) - q.x, double(ed.sy) - q.y, double(ed.sz) - q.z};
            double3 const edd{double(ed.ex) - q.x, double(ed.ey) - q.y, double(ed.ez) - q.z};
            if constexpr (Grad) {
                double3 const g = area_gradient_f64(md, sd, edd);
                gd[0] += double(ed.mult) * g.x;
                gd[1] += double(ed.mult) * g.y;
                gd[2] += double(ed.mult) * g.z;
            } else {
                double area = spherical_area_f64(md, sd, edd);
                if (axis != k_axis_z && sg > 0 && area < 0.0)
                    area += 2.0 * k_pi_d;
                sum_d += double(ed.mult) * area;
            }
        }
    }
};


// Boundary row k as the batch stages it (endpoints from the geometry, multiplicity as float).
__device__ inline staged_edge load_edge(boundary_view const &bv, std::size_t k) {
#pragma clang fp contract(off)
    std::uint32_t const a = bv.start[k], b = bv.end[k];
    return staged_edge{bv.vx[a], bv.vy[a], bv.vz[a], bv.vx[b], bv.vy[b], bv.vz[b], static_cast<float>(bv.mult[k])};
}

// One query's antipodal winding number without block cooperation (the batch kernel stages the
// chain through LDS for a whole block; this reads it from global memory): axes +Z, +X, +Y in turn
// until one is regular, crossings by BVH traversal, then the boundary edges in chain order --
// the batch's per-lane sequence of additions, so the result equals the batch's.  NaN when every
// axis is singular.  `overflowed` is set when a crossing traversal exceeded the stack (the result
// is then NaN).
template <int Prec, int Cap>
__device__ inline float antipodal_point(crossing_scene const &sc, boundary_view const &bv, v3 const &q,
                                        bool &overflowed) {
#pragma clang fp contract(off)
    overflowed = false;
    int const axes[3] = {k_axis_z, k_axis_x, k_axis_y};
    for (int ai = 0; ai < 3; ++ai) {
        int const axis = axes[ai];
        float crossings = 0.0f;
        if (crossing_count<Cap>(sc, q, axis, crossings))
            continue; // singular on this axis
        if (crossings != crossings)
            overflowed = true;
        boundary_acc<Prec, false> acc;
        for (std::size_t k = 0; k < bv.count && !acc.singular; ++k)
            acc.add(load_edge(bv, k), q, axis);
        if (acc.singular)
            continue;
        if constexpr (Prec == GWN_PRECISION_PLAIN)
            return crossings + acc.sum / (4.0f * k_pi_f);
        else
            return static_cast<float>(double(crossings) + acc.sum_d / (4.0 * k_pi_d));
    }
    return __builtin_nanf("");
}

// One query's antipodal gradient (the boundary term): (NaN, NaN, NaN) when every axis is singular.
template <int Prec>
__device__ inline void antipodal_gradient_point(boundary_view const &bv, v3 const &q, float (&g)[3]) {
#pragma clang fp contract(off)
    float const nan = __builtin_nanf("");
    g[0] = g[1] = g[2] = nan;
    int const axes[3] = {k_axis_z, k_axis_x, k_axis_y};
    for (int ai = 0; ai < 3; ++ai) {
        boundary_acc<Prec, true> acc;
        for (std::size_t k = 0; k < bv.count && !acc.singular; ++k)
            acc.add(load_edge(bv, k), q, axes[ai]);
        if (acc.singular)
            continue;
        for (int k = 0; k < 3; ++k) {
            if constexpr (Prec == GWN_PRECISION_PLAIN)
                g[k] = acc.gs[k] / (4.0f * k_pi_f);
            else
                g[k] = static_cast<float>(acc.gd[k] / (4.0 * k_pi_d));
        }
        return;
    }
}

} // namespace gwn::detail::antipodal
