// gwn_moment_math.hpp - Taylor moment payload of one triangle and the recentred node fields of one
// child slot, for any Real (float for the library's moment objects, double for Real = double).
//
// Reference: include/gwn/detail/gwn_bvh_refit_common.cuh:14-111, 124-387, 445-609.  Every function
// body is unfused (`#pragma clang fp contract(off)`), the float instantiation rounding exactly as the
// CPU oracle does.
#pragma once

#include "gwn_core.hpp"

#include <cmath>

namespace gwn {
namespace moment_math {

template <class R, int Order> struct payload_t {
    static constexpr int count = Order == 0 ? 7 : (Order == 1 ? 16 : 34);
    R v[count];
};
template <int Order> using payload = payload_t<float, Order>;

template <class R> __device__ inline R fmax_ref(R a, R b) { return (a < b) ? b : a; } // std::max(a, b)

template <class R, int Order> __device__ inline void payload_zero(payload_t<R, Order> &p) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < payload_t<R, Order>::count; ++i)
        p.v[i] = R(0);
}

template <class R, int Order> __device__ inline void payload_add(payload_t<R, Order> &d, payload_t<R, Order> const &s) {
#pragma clang fp contract(off)
#pragma unroll
    for (int i = 0; i < payload_t<R, Order>::count; ++i)
        d.v[i] += s.v[i];
}

// Second-moment integrals of one triangle along axis i (HDK's sorted-vertex split,
// gwn_bvh_refit_common.cuh:227-281).  The vertices are ordered along axis i as lo, mid, hi; the
// line through `mid` perpendicular to axis i cuts the triangle into a lower piece (lo side) and an
// upper piece (hi side), each integrated in closed form about the triangle centroid P.  Writes the
// (i, i) integral and, when requested, the (i, j) and (i, k) cross integrals.
template <class R>
__device__ inline void axis_integrals(
    R const (&verts)[3][3], R const (&P)[3], int lo, int mid, int hi, R *out_ii, R *out_ij, R *out_ik, int i
) {
#pragma clang fp contract(off)
    int const j = (i == 2) ? 0 : (i + 1);
    int const k = (j == 2) ? 0 : (j + 1);
    R const lm_i = verts[mid][i] - verts[lo][i];
    R const lm_j = verts[mid][j] - verts[lo][j];
    R const lm_k = verts[mid][k] - verts[lo][k];
    R const lh_i = verts[hi][i] - verts[lo][i];
    R const lh_j = verts[hi][j] - verts[lo][j];
    R const lh_k = verts[hi][k] - verts[lo][k];
    R const hm_i = verts[mid][i] - verts[hi][i];
    R const hm_j = verts[mid][j] - verts[hi][j];
    R const hm_k = verts[mid][k] - verts[hi][k];
    R const frac = (lh_i > R(0.0)) ? (lm_i / lh_i) : R(0.0);
    R const cut_j = frac * lh_j - lm_j;
    R const cut_k = frac * lh_k - lm_k;
    R const wl0 = cut_j * lm_k - cut_k * lm_j;
    R const wl1 = cut_k * lm_i;
    R const wl2 = cut_j * lm_i;
    R const wu0 = cut_j * hm_k - cut_k * hm_j;
    R const wu1 = cut_k * hm_i;
    R const wu2 = cut_j * hm_i;
    R const area_lo = sqrt(fmax_ref(wl0 * wl0 + wl1 * wl1 + wl2 * wl2, R(0.0)));
    R const area_hi = sqrt(fmax_ref(wu0 * wu0 + wu1 * wu1 + wu2 * wu2, R(0.0)));
    R const rlo_i = verts[lo][i] - P[i];
    R const rhi_i = verts[hi][i] - P[i];
    *out_ii = area_lo * (R(0.5) * rlo_i * rlo_i + R(2.0 / 3.0) * rlo_i * lm_i + R(0.25) * lm_i * lm_i) +
              area_hi * (R(0.5) * rhi_i * rhi_i + R(2.0 / 3.0) * rhi_i * hm_i + R(0.25) * hm_i * hm_i);
    for (int pass = 0; pass < 2; ++pass) {
        R *out = pass == 0 ? out_ij : out_ik;
        int const jk = pass == 0 ? j : k;
        R const cut = pass == 0 ? cut_j : cut_k;
        if (out == nullptr)
            continue;
        R const split = verts[mid][jk] + R(0.5) * cut;
        R const lo_split = split - verts[lo][jk];
        R const hi_split = split - verts[hi][jk];
        R const rlo_j = verts[lo][jk] - P[jk];
        R const rhi_j = verts[hi][jk] - P[jk];
        *out = area_lo * (R(0.5) * rlo_i * rlo_j + R(1.0 / 3.0) * rlo_i * lo_split +
                          R(1.0 / 3.0) * rlo_j * lm_i + R(0.25) * lm_i * lo_split) +
               area_hi * (R(0.5) * rhi_i * rhi_j + R(1.0 / 3.0) * rhi_i * hi_split +
                          R(1.0 / 3.0) * rhi_j * hm_i + R(0.25) * hm_i * hi_split);
    }
}

// Additive moments of one leaf record (gwn_bvh_refit_common.cuh:122-387), on the reference's
// reconstructed vertices b = v0 + (v1 - v0), c = v0 + (v2 - v0).
template <int Order, class R, class Record>
__device__ inline void triangle_payload(Record const &r, payload_t<R, Order> &m) {
#pragma clang fp contract(off)
    R a[3], b[3], c[3];
    {
        R const v0[3] = {r.v0x, r.v0y, r.v0z};
        R const v1[3] = {r.v1x, r.v1y, r.v1z};
        R const v2[3] = {r.v2x, r.v2y, r.v2z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            R const e1 = v1[k] - v0[k];
            R const e2 = v2[k] - v0[k];
            a[k] = v0[k];
            b[k] = v0[k] + e1;
            c[k] = v0[k] + e2;
        }
    }
    payload_zero(m);
    R const avx = (a[0] + b[0] + c[0]) / R(3.0);
    R const avy = (a[1] + b[1] + c[1]) / R(3.0);
    R const avz = (a[2] + b[2] + c[2]) / R(3.0);
    R const abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
    R const acx = c[0] - a[0], acy = c[1] - a[1], acz = c[2] - a[2];
    R const nx = R(0.5) * (aby * acz - abz * acy);
    R const ny = R(0.5) * (abz * acx - abx * acz);
    R const nz = R(0.5) * (abx * acy - aby * acx);
    R const area2 = nx * nx + ny * ny + nz * nz;
    R const area = sqrt(fmax_ref(area2, R(0.0)));
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
        if (area == R(0.0))
            return;
        R const inv_area = R(1.0) / area;
        R const nh[3] = {nx * inv_area, ny * inv_area, nz * inv_area};
        R const verts[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
        R const P[3] = {avx, avy, avz};
        int ord[3][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}};
        R span[3];
        for (int axis = 0; axis < 3; ++axis) {
            int t;
            if (verts[0][axis] > verts[1][axis]) { t = ord[axis][0]; ord[axis][0] = ord[axis][1]; ord[axis][1] = t; }
            if (verts[ord[axis][0]][axis] > verts[2][axis]) { t = ord[axis][0]; ord[axis][0] = ord[axis][2]; ord[axis][2] = t; }
            if (verts[ord[axis][1]][axis] > verts[ord[axis][2]][axis]) { t = ord[axis][1]; ord[axis][1] = ord[axis][2]; ord[axis][2] = t; }
            span[axis] = verts[ord[axis][2]][axis] - verts[ord[axis][0]][axis];
        }
        R ixx = R(0.0), iyy = R(0.0), izz = R(0.0), ixy = R(0.0), iyz = R(0.0), izx = R(0.0);
        R const dx = span[0], dy = span[1], dz = span[2];
        if (dx > R(0.0))
            axis_integrals(verts, P, ord[0][0], ord[0][1], ord[0][2], &ixx,
                           (dx >= dy && dy > R(0.0)) ? &ixy : nullptr,
                           (dx >= dz && dz > R(0.0)) ? &izx : nullptr, 0);
        if (dy > R(0.0))
            axis_integrals(verts, P, ord[1][0], ord[1][1], ord[1][2], &iyy,
                           (dy >= dz && dz > R(0.0)) ? &iyz : nullptr,
                           (dx < dy && dx > R(0.0)) ? &ixy : nullptr, 1);
        if (dz > R(0.0))
            axis_integrals(verts, P, ord[2][0], ord[2][1], ord[2][2], &izz,
                           (dx < dz && dx > R(0.0)) ? &izx : nullptr,
                           (dy < dz && dy > R(0.0)) ? &iyz : nullptr, 2);
        R const N[3] = {nx, ny, nz};
        R const I[6] = {ixx, iyy, izz, ixy, iyz, izx};
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            R *rr = m.v + 16 + 6 * comp;
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
template <int Order, class R, class Box>
__device__ inline void write_fields(R *fields, int slot, Box const &bnd, payload_t<R, Order> const &p) {
#pragma clang fp contract(off)
#define GWN_FIELD(f) fields[(f) * 4 + slot]
    R ax, ay, az;
    R const area = p.v[0];
    if (area > R(0.0)) {
        ax = p.v[1] / area;
        ay = p.v[2] / area;
        az = p.v[3] / area;
    } else {
        ax = (bnd.min_x + bnd.max_x) * R(0.5);
        ay = (bnd.min_y + bnd.max_y) * R(0.5);
        az = (bnd.min_z + bnd.max_z) * R(0.5);
    }
    R const dx = fmax_ref(ax - bnd.min_x, bnd.max_x - ax);
    R const dy = fmax_ref(ay - bnd.min_y, bnd.max_y - ay);
    R const dz = fmax_ref(az - bnd.min_z, bnd.max_z - az);
    R const Nx = p.v[4], Ny = p.v[5], Nz = p.v[6];
    GWN_FIELD(0) = dx * dx + dy * dy + dz * dz;
    GWN_FIELD(1) = ax;
    GWN_FIELD(2) = ay;
    GWN_FIELD(3) = az;
    GWN_FIELD(4) = Nx;
    GWN_FIELD(5) = Ny;
    GWN_FIELD(6) = Nz;
    if constexpr (Order >= 1) {
        R const *f = p.v + 7;
        R const nxx = f[0] - ax * Nx;
        R const nyy = f[4] - ay * Ny;
        R const nzz = f[8] - az * Nz;
        R const nxy = f[1] - ay * Nx;
        R const nyx = f[3] - ax * Ny;
        R const nyz = f[5] - az * Ny;
        R const nzy = f[7] - ay * Nz;
        R const nzx = f[6] - ax * Nz;
        R const nxz = f[2] - az * Nx;
        GWN_FIELD(7) = nxx;
        GWN_FIELD(8) = nyy;
        GWN_FIELD(9) = nzz;
        GWN_FIELD(10) = nxy + nyx;
        GWN_FIELD(11) = nyz + nzy;
        GWN_FIELD(12) = nzx + nxz;
    }
    if constexpr (Order >= 2) {
        R const Px = ax, Py = ay, Pz = az;
        R const Nc[3] = {Nx, Ny, Nz};
        R c[3][6];
#pragma unroll
        for (int comp = 0; comp < 3; ++comp) {
            R const *raw = p.v + 16 + 6 * comp;
            R const *F = p.v + 7 + 3 * comp;
            R const Nn = Nc[comp];
            c[comp][0] = raw[0] - R(2.0) * Px * F[0] + Px * Px * Nn;
            c[comp][1] = raw[1] - R(2.0) * Py * F[1] + Py * Py * Nn;
            c[comp][2] = raw[2] - R(2.0) * Pz * F[2] + Pz * Pz * Nn;
            c[comp][3] = raw[3] - Px * F[1] - Py * F[0] + Px * Py * Nn;
            c[comp][4] = raw[4] - Py * F[2] - Pz * F[1] + Py * Pz * Nn;
            c[comp][5] = raw[5] - Pz * F[0] - Px * F[2] + Pz * Px * Nn;
        }
        GWN_FIELD(13) = c[0][0];
        GWN_FIELD(14) = c[1][1];
        GWN_FIELD(15) = c[2][2];
        GWN_FIELD(16) = R(2.0) * (c[0][4] + c[1][5] + c[2][3]);
        GWN_FIELD(17) = R(2.0) * c[0][3] + c[1][0];
        GWN_FIELD(18) = R(2.0) * c[0][5] + c[2][0];
        GWN_FIELD(19) = R(2.0) * c[1][4] + c[2][1];
        GWN_FIELD(20) = R(2.0) * c[1][3] + c[0][1];
        GWN_FIELD(21) = R(2.0) * c[2][5] + c[0][2];
        GWN_FIELD(22) = R(2.0) * c[2][4] + c[1][2];
    }
#undef GWN_FIELD
}


} // namespace moment_math
} // namespace gwn
