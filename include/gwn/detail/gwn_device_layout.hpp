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
