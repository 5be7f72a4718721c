// gwn/detail/gwn_device_layout.hpp - device-side layout of libgwn_hip.so's objects (HIP, gfx950).
//
// Shared by the library's kernels and the public device point API (gwn/gwn_device.hpp), so a
// user kernel reads an object's device storage (through its accessor) with the same types the
// library writes it with.  Wide nodes of every width (2, 3, 4) are stored four-wide; slots past a
// node's children hold an invalid reference.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gwn {

// ---------------------------------------------------------------------------------------------
// Canonical BVH layout (kept byte-compatible with the reference's gwn_bvh_child<float> and
// gwn_bvh_node<4, float>: include/gwn/gwn_bvh.cuh:25-130).
// ---------------------------------------------------------------------------------------------

constexpr int k_offset_bits = 47;
constexpr std::uint64_t k_offset_mask = (std::uint64_t(1) << k_offset_bits) - 1;
constexpr std::uint64_t k_valid_mask = std::uint64_t(1) << k_offset_bits;
constexpr int k_count_shift = k_offset_bits + 1;

struct aabb {
    float min_x, min_y, min_z, max_x, max_y, max_z;
};

struct bvh_child {
    aabb bounds;
    std::uint64_t reference;
};
static_assert(sizeof(bvh_child) == 32, "child must be one 32-byte record");

struct bvh_node4 {
    bvh_child children[4];
};
static_assert(sizeof(bvh_node4) == 128, "BVH4 node must be 128 bytes");

__host__ __device__ constexpr std::uint64_t encode_internal(std::uint64_t node) {
    return k_valid_mask | node;
}
__host__ __device__ constexpr std::uint64_t encode_leaf(std::uint64_t offset, std::uint64_t count) {
    return k_valid_mask | offset | (count << k_count_shift);
}
__host__ __device__ constexpr bool ref_valid(std::uint64_t r) { return (r & k_valid_mask) != 0; }
__host__ __device__ constexpr std::uint32_t ref_count(std::uint64_t r) {
    return static_cast<std::uint32_t>(r >> k_count_shift);
}
__host__ __device__ constexpr bool ref_internal(std::uint64_t r) {
    return ref_valid(r) && ref_count(r) == 0;
}
__host__ __device__ constexpr bool ref_leaf(std::uint64_t r) { return ref_valid(r) && ref_count(r) != 0; }
__host__ __device__ constexpr std::uint64_t ref_offset(std::uint64_t r) { return r & k_offset_mask; }

// Compact 32-bit child reference used by the traversal records of the moment object:
// bit 31 set = leaf (sorted primitive offset in bits 0-30), otherwise an internal node index.
constexpr std::uint32_t k_cref_invalid = 0xffffffffu;
constexpr std::uint32_t k_cref_leaf = 0x80000000u;

// Leaf-ordered triangle record: the three vertices, 36 bytes.  The reference's
// (v0, e1, e2) form is recovered exactly as e1 = v1 - v0 (include/gwn/detail/
// gwn_bvh_triangle_impl.cuh:56-59), and the exact vertices feed the FP64 near field.
struct tri_record {
    float v0x, v0y, v0z, v1x, v1y, v1z, v2x, v2y, v2z;
};
static_assert(sizeof(tri_record) == 36, "triangle record must be 36 bytes");

// Number of float fields per child slot of a Taylor moment node (include/gwn/gwn_bvh.cuh:144-202).
__host__ __device__ constexpr int moment_field_count(int order) {
    return order == 0 ? 7 : (order == 1 ? 13 : 23);
}
// Traversal record: one float4 of compact child references, then the fields as float4 (slot-SoA),
// padded to a whole number of 128-byte cache lines: 128 B / 256 B / 384 B for order 0 / 1 / 2.
__host__ __device__ constexpr int moment_record_vec4(int order) {
    return order == 0 ? 8 : (order == 1 ? 16 : 24);
}

// Evaluation record: the traversal record with the expansion's wave-uniform arithmetic folded in
// at refit time, for the value-only Taylor queries.  Its first five float4 are the traversal
// record's (child references, then max_p_dist2 and the expansion centre per slot: what the
// acceptance test reads); then one block of eval_slot_floats(order) floats per child slot holding
// that slot's coefficients contiguously (N, then M, then C'; zero padded), so an accepted slot's
// coefficients arrive in one or two scalar loads.  With q the unit direction and |q| = 1 the
// order-1 and order-2 terms become plain forms in q:
//   Tr - 3 q^T T q               = q^T M q,  M = Tr I - 3 T                       (fields 7-12)
//   1.5 (q . P) - 7.5 C(q)       = C'(q) = 1.5 (q . P)(q . q) - 7.5 C(q)          (fields 13-22)
// with P = (3 xxx + Ayx + Azx, 3 yyy + Azy + Axy, 3 zzz + Axz + Ayz) and C the reference's cubic
// (gwn_query_winding_impl.cuh:186-325).  Field 7, 8, 9 hold M's diagonal, 10, 11, 12 the full
// coefficients of qx qy, qy qz, qz qx; fields 13-22 the coefficients of qx^3, qy^3, qz^3,
// qx qy qz, qx^2 qy, qx^2 qz, qy^2 qz, qy^2 qx, qz^2 qx, qz^2 qy.  `c` is one slot's traversal
// fields (c[f] = field f), `e` receives its evaluation fields in the same numbering; a slot's
// block holds e[4], e[5], ..., e[F - 1].
template <int Order>
__host__ __device__ inline void taylor_eval_fields(float const *c, float *e) {
#pragma clang fp contract(off)
    constexpr int F = moment_field_count(Order);
    for (int f = 0; f < 7; ++f)
        e[f] = c[f];
    if constexpr (Order >= 1) {
        float const tr = (c[7] + c[8]) + c[9];
        e[7] = tr - 3.0f * c[7];
        e[8] = tr - 3.0f * c[8];
        e[9] = tr - 3.0f * c[9];
        e[10] = -3.0f * c[10];
        e[11] = -3.0f * c[11];
        e[12] = -3.0f * c[12];
    }
    if constexpr (Order >= 2) {
        float const xxx = c[13], yyy = c[14], zzz = c[15], spn = c[16];
        float const axy = c[17], axz = c[18], ayz = c[19], ayx = c[20], azx = c[21], azy = c[22];
        float const px = 1.5f * ((3.0f * xxx + ayx) + azx);
        float const py = 1.5f * ((3.0f * yyy + azy) + axy);
        float const pz = 1.5f * ((3.0f * zzz + axz) + ayz);
        e[13] = px - 7.5f * xxx;
        e[14] = py - 7.5f * yyy;
        e[15] = pz - 7.5f * zzz;
        e[16] = -7.5f * spn;
        e[17] = py - 7.5f * axy;
        e[18] = pz - 7.5f * axz;
        e[19] = pz - 7.5f * ayz;
        e[20] = px - 7.5f * ayx;
        e[21] = px - 7.5f * azx;
        e[22] = py - 7.5f * azy;
    }
    (void)F;
}

__host__ __device__ constexpr int eval_slot_floats(int order) { return order == 0 ? 4 : (order == 1 ? 12 : 20); }
// float4 per evaluation record: 5 (references, acceptance fields) + 4 slot blocks.
__host__ __device__ constexpr int eval_record_vec4(int order) { return 5 + eval_slot_floats(order); }

// Nearest-triangle codes of the unsigned-distance batch's optional per-query output: the BVH
// leaf record of the nearest triangle, or one of these.
constexpr std::uint32_t k_nearest_none = 0xffffffffu;     // the band or zero: no gradient
constexpr std::uint32_t k_nearest_overflow = 0xfffffffeu; // the traversal stack overflowed

// Work tallies of one query's traversal, for roofline accounting: child bounding-box tests and
// triangle tests (point-triangle distances or ray-triangle intersections).
struct traversal_work {
    std::uint32_t boxes = 0, tris = 0;
};

// Five-comparator sorting network: orders t ascending and applies the same swaps to k.  Used
// for near-first child ordering in traversals.
template <class Key>
__device__ inline void sort4(float (&t)[4], Key (&k)[4]) {
    auto cs = [&](int a, int b) {
        if (!(t[a] > t[b]))
            return;
        float const tt = t[a];
        t[a] = t[b];
        t[b] = tt;
        Key const kk = k[a];
        k[a] = k[b];
        k[b] = kk;
    };
    cs(0, 1);
    cs(2, 3);
    cs(0, 2);
    cs(1, 3);
    cs(1, 2);
}

} // namespace gwn
