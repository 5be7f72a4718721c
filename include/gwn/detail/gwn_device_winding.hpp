// gwn/detail/gwn_device_winding.hpp - per-query winding-number evaluation on gfx950: the solid
// angle of one triangle (reference form, FP64, precision split), its gradient, the Taylor far-field
// terms of orders 0-2 and the per-lane Taylor traversal.
//
// Reference: include/gwn/detail/gwn_query_winding_impl.cuh:20-374,
// include/gwn/detail/gwn_query_geometry_impl.cuh:16-45.
//
// Shared by libgwn_hip.so's winding kernels and the public device point API (gwn/gwn_device.hpp),
// so a query evaluated in a user kernel is rounded exactly as the batch rounds it.  Products are
// contracted into FMAs only within one source expression: every function body sets
// `#pragma clang fp contract(on)` for itself, so including this header does not change the
// includer's contraction mode.  Build with -fhip-fp32-correctly-rounded-divide-sqrt (as the library
// is) for identical float division and square roots.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../gwn_hip.h"
#include "gwn_device_layout.hpp"

namespace gwn::detail::winding {

constexpr float k_pi_f = 3.14159265358979323846f;
constexpr double k_pi_d = 3.14159265358979323846;

// Reference-form FP32 solid angle: rebuild b = v0 + (v1 - v0), c = v0 + (v2 - v0) as the
// reference's (v0, e1, e2) record does, then Van Oosterom-Strackee.
__device__ inline float solid_angle_plain(tri_record const &t, float qx, float qy, float qz) {
#pragma clang fp contract(on)
    float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    float ax = t.v0x - qx, ay = t.v0y - qy, az = t.v0z - qz;
    float bx = (t.v0x + e1x) - qx, by = (t.v0y + e1y) - qy, bz = (t.v0z + e1z) - qz;
    float cx = (t.v0x + e2x) - qx, cy = (t.v0y + e2y) - qy, cz = (t.v0z + e2z) - qz;
    float const al = sqrtf(ax * ax + ay * ay + az * az);
    float const bl = sqrtf(bx * bx + by * by + bz * bz);
    float const cl = sqrtf(cx * cx + cy * cy + cz * cz);
    if (al == 0.0f || bl == 0.0f || cl == 0.0f)
        return 0.0f;
    ax /= al; ay /= al; az /= al;
    bx /= bl; by /= bl; bz /= bl;
    cx /= cl; cy /= cl; cz /= cl;
    float const ux = bx - ax, uy = by - ay, uz = bz - az;
    float const vx = cx - ax, vy = cy - ay, vz = cz - az;
    float const num = ax * (uy * vz - uz * vy) + ay * (uz * vx - ux * vz) + az * (ux * vy - uy * vx);
    if (num == 0.0f)
        return 0.0f;
    float const den = 1.0f + (ax * bx + ay * by + az * bz) + (ax * cx + ay * cy + az * cz) +
                      (bx * cx + by * cy + bz * cz);
    return 2.0f * atan2f(num, den);
}

// Solid angle without normalizing the corner vectors: with a = v0 - q, b = v1 - q, c = v2 - q,
// tan(omega / 2) = a . (e1 x e2) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|), which equals the
// reference's normalized form (divide numerator and denominator by |a||b||c|; qa . ((qb - qa) x
// (qc - qa)) = qa . (qb x qc)).  The numerator uses the edges e1 = v1 - v0, e2 = v2 - v0, which do
// not depend on q, so a far triangle's nearly parallel corner vectors never cancel.
template <class R>
__device__ inline R solid_angle_edges(R ax, R ay, R az, R e1x, R e1y, R e1z, R e2x, R e2y, R e2z) {
#pragma clang fp contract(on)
    R const bx = ax + e1x, by = ay + e1y, bz = az + e1z;
    R const cx = ax + e2x, cy = ay + e2y, cz = az + e2z;
    R const la = sqrt(ax * ax + ay * ay + az * az);
    R const lb = sqrt(bx * bx + by * by + bz * bz);
    R const lc = sqrt(cx * cx + cy * cy + cz * cz);
    if (la == R(0) || lb == R(0) || lc == R(0))
        return R(0);
    R const nx = e1y * e2z - e1z * e2y;
    R const ny = e1z * e2x - e1x * e2z;
    R const nz = e1x * e2y - e1y * e2x;
    R const num = ax * nx + ay * ny + az * nz;
    if (num == R(0))
        return R(0);
    R const den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                  (bx * cx + by * cy + bz * cz) * la;
    return R(2) * atan2(num, den);
}

// FP64 solid angle on the exact vertices (float inputs convert exactly, so e1 and e2 are exact).
__device__ inline double solid_angle_f64(tri_record const &t, double qx, double qy, double qz) {
#pragma clang fp contract(on)
    return solid_angle_edges<double>(
        double(t.v0x) - qx, double(t.v0y) - qy, double(t.v0z) - qz, double(t.v1x) - double(t.v0x),
        double(t.v1y) - double(t.v0y), double(t.v1z) - double(t.v0z), double(t.v2x) - double(t.v0x),
        double(t.v2y) - double(t.v0y), double(t.v2z) - double(t.v0z)
    );
}

// FP64 square root for the split near field: hardware reciprocal square root, then two Newton
// steps on the root (g = x r, h = r / 2; g += (x - g g) h, h += (1/2 - g h) h); relative error at
// FP64 rounding level, not correctly rounded.  x > 0.
__device__ inline double sqrt_f64_fast(double x) {
#pragma clang fp contract(on)
    double const r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double const d = fma(-g, g, x);
    return fma(d, h, g);
}

// Precision-split near triangle: corner vectors (exact differences of the float vertices and
// query), lengths, numerator and denominator in FP64, where the cancellation happens; the final
// angle by FP32 atan2 of the FP64 pair scaled by a power of two into range.  The angle's absolute
// error is a few 1e-7 rad, 1e-8 of W, per triangle.
__device__ inline double solid_angle_split_near(tri_record const &t, float x, float y, float z) {
#pragma clang fp contract(on)
    double const qx = x, qy = y, qz = z;
    double const ax = double(t.v0x) - qx, ay = double(t.v0y) - qy, az = double(t.v0z) - qz;
    double const e1x = double(t.v1x) - double(t.v0x), e1y = double(t.v1y) - double(t.v0y),
                 e1z = double(t.v1z) - double(t.v0z);
    double const e2x = double(t.v2x) - double(t.v0x), e2y = double(t.v2y) - double(t.v0y),
                 e2z = double(t.v2z) - double(t.v0z);
    double const bx = ax + e1x, by = ay + e1y, bz = az + e1z;
    double const cx = ax + e2x, cy = ay + e2y, cz = az + e2z;
    double const aa = ax * ax + ay * ay + az * az;
    double const bb = bx * bx + by * by + bz * bz;
    double const cc = cx * cx + cy * cy + cz * cz;
    if (aa == 0.0 || bb == 0.0 || cc == 0.0)
        return 0.0;
    double const la = sqrt_f64_fast(aa), lb = sqrt_f64_fast(bb), lc = sqrt_f64_fast(cc);
    double const nx = e1y * e2z - e1z * e2y;
    double const ny = e1z * e2x - e1x * e2z;
    double const nz = e1x * e2y - e1y * e2x;
    double const num = ax * nx + ay * ny + az * nz;
    if (num == 0.0)
        return 0.0;
    double const den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                       (bx * cx + by * cy + bz * cz) * la;
    int const e = __builtin_amdgcn_frexp_exp(fmax(fabs(num), fabs(den)));
    float const fn = static_cast<float>(__builtin_amdgcn_ldexp(num, -e));
    float const fd = static_cast<float>(__builtin_amdgcn_ldexp(den, -e));
    return 2.0 * static_cast<double>(atan2f(fn, fd));
}

// FP32 far-field solid angle (precision split): same form with hardware square roots.
__device__ inline float solid_angle_far_f32(float ax, float ay, float az, float e1x, float e1y, float e1z,
                                            float e2x, float e2y, float e2z) {
#pragma clang fp contract(on)
    float const bx = ax + e1x, by = ay + e1y, bz = az + e1z;
    float const cx = ax + e2x, cy = ay + e2y, cz = az + e2z;
    float const la = __builtin_amdgcn_sqrtf(ax * ax + ay * ay + az * az);
    float const lb = __builtin_amdgcn_sqrtf(bx * bx + by * by + bz * bz);
    float const lc = __builtin_amdgcn_sqrtf(cx * cx + cy * cy + cz * cz);
    float const nx = e1y * e2z - e1z * e2y;
    float const ny = e1z * e2x - e1x * e2z;
    float const nz = e1x * e2y - e1y * e2x;
    float const num = ax * nx + ay * ny + az * nz;
    float const den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                      (bx * cx + by * cy + bz * cz) * la;
    return 2.0f * atan2f(num, den);
}


// Near-field radius of the precision split, in units of the triangle's longest edge from v0.
constexpr float k_split_near_scale2 = 16.0f; // (4 * edge)^2

// Block-scoped contract(off): the IR operations carry no contraction flag, so they stay
// separate multiplies and adds even after inlining into FMA-friendly code.
__device__ inline float squared_norm_unfused(float x, float y, float z) {
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

__device__ inline float product_unfused(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}

// Gradient of one triangle's solid angle for the corner vectors g_i = v_i - q (Biot-Savart form,
// gwn_query_winding_impl.cuh:20-61): the sum over the edges (i, j) = (0,1), (1,2), (2,0) of
// ((g_i - g_j) . (g_i/|g_i| - g_j/|g_j|)) / |g_i x g_j|^2 * (g_i x g_j).  Zero when q sits on a
// vertex; an edge seen end-on contributes nothing.
template <class R>
__device__ inline void solid_angle_gradient(R const (&g)[3][3], R (&out)[3]) {
#pragma clang fp contract(on)
    // Unfused: near an edge the 1 / |g_i x g_j|^2 factor amplifies rounding, and the unfused
    // sequence is the oracle's, operation for operation.
#pragma clang fp contract(off)
    R l[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        l[i] = sqrt(g[i][0] * g[i][0] + g[i][1] * g[i][1] + g[i][2] * g[i][2]);
    out[0] = out[1] = out[2] = R(0);
    if (l[0] == R(0) || l[1] == R(0) || l[2] == R(0))
        return;
    R h[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            h[i][k] = g[i][k] / l[i];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        int const i = e, j = (e + 1) % 3;
        R const cx = g[i][1] * g[j][2] - g[i][2] * g[j][1];
        R const cy = g[i][2] * g[j][0] - g[i][0] * g[j][2];
        R const cz = g[i][0] * g[j][1] - g[i][1] * g[j][0];
        R const cs = cx * cx + cy * cy + cz * cz;
        R tx = R(0), ty = R(0), tz = R(0);
        if (cs != R(0)) {
            R const sc = ((g[i][0] - g[j][0]) * (h[i][0] - h[j][0]) + (g[i][1] - g[j][1]) * (h[i][1] - h[j][1]) +
                          (g[i][2] - g[j][2]) * (h[i][2] - h[j][2])) / cs;
            tx = sc * cx;
            ty = sc * cy;
            tz = sc * cz;
        }
        if (e == 0) {
            out[0] = tx; out[1] = ty; out[2] = tz;
        } else {
            out[0] += tx; out[1] += ty; out[2] += tz;
        }
    }
}

// Reference-form FP32 gradient: corners rebuilt from (v0, e1 = v1 - v0, e2 = v2 - v0).
__device__ inline void gradient_plain(tri_record const &t, float x, float y, float z, float (&out)[3]) {
#pragma clang fp contract(on)
    float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    float const g[3][3] = {{t.v0x - x, t.v0y - y, t.v0z - z},
                           {(t.v0x + e1x) - x, (t.v0y + e1y) - y, (t.v0z + e1z) - z},
                           {(t.v0x + e2x) - x, (t.v0y + e2y) - y, (t.v0z + e2z) - z}};
    solid_angle_gradient<float>(g, out);
}

// Precision split for one leaf triangle reached by the traversal: a triangle whose first vertex
// is within four edge lengths of the query (the atan2-cancellation zone) is evaluated in FP64 on
// the exact vertices; a farther one in FP32 (the gradient from the exact vertices).  Everything
// goes to the query's FP64 sums.  The distance test is unfused so every implementation
// classifies a triangle the same way.
template <bool Grad>
__device__ inline void split_leaf(tri_record const &t, float x, float y, float z, double &sum, double (&gsum)[3]) {
#pragma clang fp contract(on)
    float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
    float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
    float const ax = t.v0x - x, ay = t.v0y - y, az = t.v0z - z;
    float const le1 = squared_norm_unfused(e1x, e1y, e1z);
    float const le2 = squared_norm_unfused(e2x, e2y, e2z);
    float const dq = squared_norm_unfused(ax, ay, az);
    if (!(dq > product_unfused(k_split_near_scale2, fmaxf(le1, le2)))) {
        sum += solid_angle_split_near(t, x, y, z);
        if constexpr (Grad) {
            double const dx = x, dy = y, dz = z;
            double const g[3][3] = {{double(t.v0x) - dx, double(t.v0y) - dy, double(t.v0z) - dz},
                                    {double(t.v1x) - dx, double(t.v1y) - dy, double(t.v1z) - dz},
                                    {double(t.v2x) - dx, double(t.v2y) - dy, double(t.v2z) - dz}};
            double o[3];
            solid_angle_gradient<double>(g, o);
            gsum[0] += o[0];
            gsum[1] += o[1];
            gsum[2] += o[2];
        }
    } else {
        sum += double(solid_angle_far_f32(ax, ay, az, e1x, e1y, e1z, e2x, e2y, e2z));
        if constexpr (Grad) {
            float const g[3][3] = {{ax, ay, az}, {t.v1x - x, t.v1y - y, t.v1z - z}, {t.v2x - x, t.v2y - y, t.v2z - z}};
            float o[3];
            solid_angle_gradient<float>(g, o);
            gsum[0] += double(o[0]);
            gsum[1] += double(o[1]);
            gsum[2] += double(o[2]);
        }
    }
}

// One accepted far-field expansion from a slot's evaluation block `blk` (blk[k] = evaluation
// field k + 4, taylor_eval_fields): the value of taylor_term below with its wave-uniform arithmetic
// folded into the record,
//   w = m2 (m1 (q^T M q + m1 C'(q)) - q . N),  m1 = 1 / |q - P|,  q the unit direction.
// Equal to the reference's expression up to rounding (and |q| = 1 within a few ulps).
template <int Order>
__device__ inline float taylor_value(float const *blk, float qrx, float qry, float qrz, float ql2) {
#pragma clang fp contract(on)
#define GWN_F(k) blk[(k) - 4]
    float const m1 = __builtin_amdgcn_rsqf(ql2);
    float const m2 = m1 * m1;
    float const qnx = qrx * m1, qny = qry * m1, qnz = qrz * m1;
    float const rdotN = qnx * GWN_F(4) + qny * GWN_F(5) + qnz * GWN_F(6);
    if constexpr (Order == 0) {
        return -m2 * rdotN;
    } else {
        float const qz2 = qnz * qnz;
        float const a = qnx * GWN_F(7) + qny * GWN_F(10) + qnz * GWN_F(12);
        float const b = qny * GWN_F(8) + qnz * GWN_F(11);
        float t = qnx * a + qny * b + qz2 * GWN_F(9);
        if constexpr (Order >= 2) {
            float const qx2 = qnx * qnx, qy2 = qny * qny;
            float const cx = qnx * GWN_F(13) + qny * GWN_F(17) + qnz * GWN_F(18);
            float const cy = qny * GWN_F(14) + qnz * GWN_F(19) + qnx * GWN_F(20);
            float const cz = qnz * GWN_F(15) + qnx * GWN_F(21) + qny * GWN_F(22);
            float const cub = qx2 * cx + qy2 * cy + qz2 * cz + (qnx * qny) * (qnz * GWN_F(16));
            t = t + m1 * cub;
        }
        return m2 * (m1 * t - rdotN);
    }
#undef GWN_F
}

// One accepted far-field expansion (gwn_query_winding_impl.cuh:186-325): returns the solid-angle
// term and, with Grad, writes its gradient.
template <int Order, bool Grad>
__device__ inline float taylor_term(float const *f, int s, float qrx, float qry, float qrz, float ql2, float (&g)[3]) {
#pragma clang fp contract(on)
#define GWN_F(k) f[(k) * 4 + s]
    float const m1 = __builtin_amdgcn_rsqf(ql2);
    float const m2 = m1 * m1;
    float const m3 = m2 * m1;
    float const qnx = qrx * m1, qny = qry * m1, qnz = qrz * m1;
    float const Nx = GWN_F(4), Ny = GWN_F(5), Nz = GWN_F(6);
    float const rdotN = qnx * Nx + qny * Ny + qnz * Nz;
    float w = -m2 * rdotN;
    if constexpr (Grad) {
        g[0] = m3 * (3.0f * qnx * rdotN - Nx);
        g[1] = m3 * (3.0f * qny * rdotN - Ny);
        g[2] = m3 * (3.0f * qnz * rdotN - Nz);
    }
    if constexpr (Order >= 1) {
        float const qxx = qnx * qnx, qyy = qny * qny, qzz = qnz * qnz;
        float const Txx = GWN_F(7), Tyy = GWN_F(8), Tzz = GWN_F(9);
        float const Sxy = GWN_F(10), Syz = GWN_F(11), Szx = GWN_F(12);
        float const Tr = Txx + Tyy + Tzz;
        float const S = qxx * Txx + qyy * Tyy + qzz * Tzz + qnx * qny * Sxy + qnx * qnz * Szx + qny * qnz * Syz;
        w += m3 * (Tr - 3.0f * S);
        if constexpr (Grad) {
            float const m4 = m2 * m2;
            float const vx = 2.0f * Txx * qnx + Sxy * qny + Szx * qnz;
            float const vy = Sxy * qnx + 2.0f * Tyy * qny + Syz * qnz;
            float const vz = Szx * qnx + Syz * qny + 2.0f * Tzz * qnz;
            float const coeff = 15.0f * S - 3.0f * Tr;
            g[0] += m4 * (coeff * qnx - 3.0f * vx);
            g[1] += m4 * (coeff * qny - 3.0f * vy);
            g[2] += m4 * (coeff * qnz - 3.0f * vz);
        }
    }
    if constexpr (Order >= 2) {
        float const m4 = m2 * m2;
        float const qnx2 = qnx * qnx, qny2 = qny * qny, qnz2 = qnz * qnz;
        float const qnx3 = qnx2 * qnx, qny3 = qny2 * qny, qnz3 = qnz2 * qnz;
        float const xxx = GWN_F(13), yyy = GWN_F(14), zzz = GWN_F(15), spn = GWN_F(16);
        float const Axy = GWN_F(17), Axz = GWN_F(18), Ayz = GWN_F(19), Ayx = GWN_F(20), Azx = GWN_F(21), Azy = GWN_F(22);
        float const t0x = Ayx + Azx, t0y = Azy + Axy, t0z = Axz + Ayz;
        float const t1x = qny * Axy + qnz * Axz;
        float const t1y = qnz * Ayz + qnx * Ayx;
        float const t1z = qnx * Azx + qny * Azy;
        w += m4 * (1.5f * (qnx * (3.0f * xxx + t0x) + qny * (3.0f * yyy + t0y) + qnz * (3.0f * zzz + t0z)) -
                   7.5f * (qnx3 * xxx + qny3 * yyy + qnz3 * zzz + qnx * qny * qnz * spn + qnx2 * t1x +
                           qny2 * t1y + qnz2 * t1z));
        if constexpr (Grad) {
            float const m5 = m4 * m1;
            float const Px = 3.0f * xxx + Ayx + Azx;
            float const Py = 3.0f * yyy + Azy + Axy;
            float const Pz = 3.0f * zzz + Axz + Ayz;
            float const L2 = qnx * Px + qny * Py + qnz * Pz;
            float const C2 = qnx3 * xxx + qny3 * yyy + qnz3 * zzz + qnx * qny * qnz * spn +
                             qnx2 * (qny * Axy + qnz * Axz) + qny2 * (qnz * Ayz + qnx * Ayx) +
                             qnz2 * (qnx * Azx + qny * Azy);
            float const Rx = 3.0f * xxx * qnx2 + 2.0f * Axy * qnx * qny + 2.0f * Axz * qnx * qnz + Ayx * qny2 +
                             Azx * qnz2 + spn * qny * qnz;
            float const Ry = 3.0f * yyy * qny2 + 2.0f * Ayx * qnx * qny + 2.0f * Ayz * qny * qnz + Axy * qnx2 +
                             Azy * qnz2 + spn * qnx * qnz;
            float const Rz = 3.0f * zzz * qnz2 + 2.0f * Azx * qnx * qnz + 2.0f * Azy * qny * qnz + Axz * qnx2 +
                             Ayz * qny2 + spn * qnx * qny;
            float const h = 52.5f * C2;
            g[0] += m5 * (1.5f * Px - 7.5f * (qnx * L2 + Rx) + qnx * h);
            g[1] += m5 * (1.5f * Py - 7.5f * (qny * L2 + Ry) + qny * h);
            g[2] += m5 * (1.5f * Pz - 7.5f * (qnz * L2 + Rz) + qnz * h);
        }
    }
#undef GWN_F
    return w;
}

constexpr float k_inv_4pi_f = 1.0f / (4.0f * k_pi_f);

// Per-query sums of one traversal: the reference's float sums (plain) or FP64 sums (split).
template <int Prec, bool Grad> struct taylor_sums {
    float omega = 0.0f;
    float gf[3] = {0.0f, 0.0f, 0.0f};
    double near = 0.0;
    double gd[3] = {0.0, 0.0, 0.0};

    __device__ void far_term(float w, float const (&g)[3]) {
#pragma clang fp contract(on)
        if constexpr (Prec == GWN_PRECISION_PLAIN) {
            omega += w;
            if constexpr (Grad) {
                gf[0] += g[0];
                gf[1] += g[1];
                gf[2] += g[2];
            }
        } else {
            near += double(w);
            if constexpr (Grad) {
                gd[0] += double(g[0]);
                gd[1] += double(g[1]);
                gd[2] += double(g[2]);
            }
        }
    }
    __device__ void leaf(tri_record const &t, float x, float y, float z) {
#pragma clang fp contract(on)
        if constexpr (Prec == GWN_PRECISION_PLAIN) {
            omega += solid_angle_plain(t, x, y, z);
            if constexpr (Grad) {
                float o[3];
                gradient_plain(t, x, y, z, o);
                gf[0] += o[0];
                gf[1] += o[1];
                gf[2] += o[2];
            }
        } else {
            split_leaf<Grad>(t, x, y, z, near, gd);
        }
    }
    // Winding = sum / 4 pi; gradient = sum * (1 / 4 pi) (gwn_query_winding_impl.cuh:130-139).
    __device__ void store(bool overflow, std::size_t q, float *out, float *gx, float *gy, float *gz) const {
#pragma clang fp contract(on)
        float const nan = __builtin_nanf("");
        float w;
        float g[3];
        if (overflow) {
            w = g[0] = g[1] = g[2] = nan;
        } else if constexpr (Prec == GWN_PRECISION_PLAIN) {
            w = omega / (4.0f * k_pi_f);
            g[0] = gf[0] * k_inv_4pi_f;
            g[1] = gf[1] * k_inv_4pi_f;
            g[2] = gf[2] * k_inv_4pi_f;
        } else {
            w = static_cast<float>((double(omega) + near) / (4.0 * k_pi_d));
            double const inv = 1.0 / (4.0 * k_pi_d);
            g[0] = static_cast<float>(gd[0] * inv);
            g[1] = static_cast<float>(gd[1] * inv);
            g[2] = static_cast<float>(gd[2] * inv);
        }
        if (out != nullptr)
            out[q] = w;
        if constexpr (Grad) {
            gx[q] = g[0];
            gy[q] = g[1];
            gz[q] = g[2];
        }
    }
};

struct taylor_outputs {
    float *out;                 // winding numbers (may be null when only the gradient is asked)
    float *gx, *gy, *gz;        // gradient (used when the kernel computes it)
    std::uint32_t *counters;    // optional 3 x uint32 per query
};

// Per-lane Taylor traversal of one query from an internal root: children in slot order, the last
// descended internal child entered directly, the others pushed.  Returns true on stack overflow.
// `records` are the moments' evaluation records for the value alone (Grad = false) and their
// traversal records with the gradient (Grad = true).
template <int Order, int Prec, bool Grad, int StackCap>
__device__ inline bool taylor_traverse(
    float4 const *__restrict__ records, tri_record const *__restrict__ tris, std::uint32_t root_node, float x,
    float y, float z, float beta2, taylor_sums<Prec, Grad> &acc, std::uint32_t &n_tests, std::uint32_t &n_accept,
    std::uint32_t &n_leaf
) {
#pragma clang fp contract(on)
    constexpr int RV = Grad ? moment_record_vec4(Order) : eval_record_vec4(Order);
    constexpr int B = eval_slot_floats(Order);
    std::uint32_t stack[StackCap];
    int sp = 0;
    bool overflow = false;
    std::uint32_t node = root_node;
    // One loop exit, at the head: an early return from inside the unrolled child loop has been
    // lowered wrongly by the gfx950 backend (ROCm 7.2 LLVM) in the same traversal shape.
    while (node != k_cref_invalid && !overflow) {
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
        std::uint32_t next = k_cref_invalid;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            std::uint32_t const c = cref[s];
            if (c == k_cref_invalid || overflow)
                continue;
            ++n_tests;
            float const qrx = x - cx[s], qry = y - cy[s], qrz = z - cz[s];
            // The acceptance test is evaluated unfused, operation by operation as the reference
            // writes it, so every implementation takes the same accept/descend decisions.
            float const ql2 = squared_norm_unfused(qrx, qry, qrz);
            bool const descend = !(ql2 > 0.0f) || ql2 <= product_unfused(mpd[s], beta2);
            if (!descend) {
                float g[3];
                float w;
                if constexpr (Grad)
                    w = taylor_term<Order, true>(fields, s, qrx, qry, qrz, ql2, g);
                else
                    w = taylor_value<Order>(reinterpret_cast<float const *>(rec + 5) + s * B, qrx, qry, qrz, ql2);
                if (isfinite(w)) {
                    acc.far_term(w, g);
                    ++n_accept;
                    continue;
                }
            }
            if (c & k_cref_leaf) {
                ++n_leaf;
                acc.leaf(tris[c & ~k_cref_leaf], x, y, z);
                continue;
            }
            if (next != k_cref_invalid) {
                if (sp < StackCap)
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

} // namespace gwn::detail::winding
