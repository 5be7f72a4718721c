// gwn_distance.hip - unsigned point-to-mesh distance batch on gfx950.
//
// Reference: include/gwn/detail/gwn_query_distance_impl.cuh:20-192,
// include/gwn/detail/gwn_query_geometry_impl.cuh:46-114, include/gwn/gwn_query.cuh:109-180.
//
// One query per lane, near-first culled traversal of the canonical BVH4: the culling band
// (clamped at zero) is the initial squared bound and the result when nothing is closer; children
// whose box is no closer than the current best are skipped; the survivors are ordered by box
// distance with a five-comparator network; leaves of a node are evaluated before its internal
// children are queued, the nearest internal child is entered directly and the others are pushed
// far-to-near.  A query that reaches distance zero stops (zero is the global lower bound).
// Queries may be visited through a Morton-order permutation so that a wave's lanes walk nearby
// paths; results land at the input positions.
//
// Arithmetic is unfused and the square root correctly rounded, so every result equals the CPU
// oracle's replay bit for bit.
#include "gwn_core.hpp"

#pragma clang fp contract(off)

namespace gwn {
namespace {

constexpr int k_block = 256;

__device__ inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

// Squared distance from p to the closed triangle (a, b, c) by Voronoi regions (Ericson).
__device__ inline float point_triangle_distance2(
    float px, float py, float pz, float ax, float ay, float az, float bx, float by, float bz, float cx,
    float cy, float cz
) {
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
    float const dx = max0(b.min_x - qx) + max0(qx - b.max_x);
    float const dy = max0(b.min_y - qy) + max0(qy - b.max_y);
    float const dz = max0(b.min_z - qz) + max0(qz - b.max_z);
    return dx * dx + dy * dy + dz * dz;
}

// Leaf range: records rebuilt in the reference's (v0, v0 + e1, v0 + e2) form.
__device__ inline bool visit_leaf(
    tri_record const *__restrict__ tris, std::size_t nt, std::uint64_t ref, float qx, float qy, float qz,
    float &best
) {
    std::uint64_t const off = ref_offset(ref);
    std::uint32_t const cnt = ref_count(ref);
    for (std::uint32_t k = 0; k < cnt; ++k) {
        if (off + k >= nt)
            continue;
        tri_record const t = tris[off + k];
        float const e1x = t.v1x - t.v0x, e1y = t.v1y - t.v0y, e1z = t.v1z - t.v0z;
        float const e2x = t.v2x - t.v0x, e2y = t.v2y - t.v0y, e2z = t.v2z - t.v0z;
        float const d2 = point_triangle_distance2(qx, qy, qz, t.v0x, t.v0y, t.v0z, t.v0x + e1x, t.v0y + e1y,
                                                  t.v0z + e1z, t.v0x + e2x, t.v0y + e2y, t.v0z + e2z);
        if (d2 < best)
            best = d2;
    }
    return best > 0.0f;
}

template <int StackCap>
__global__ __launch_bounds__(k_block) void distance_kernel(
    bvh_node4 const *__restrict__ nodes, std::size_t node_count, tri_record const *__restrict__ tris,
    std::size_t nt, std::uint64_t root_ref, float const *__restrict__ qx, float const *__restrict__ qy,
    float const *__restrict__ qz, std::uint32_t const *__restrict__ order, float *__restrict__ out,
    std::size_t n, float culling_band
) {
    std::size_t const i = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (i >= n)
        return;
    std::size_t const q = order ? order[i] : i;
    float const x = qx[q], y = qy[q], z = qz[q];
    float const band = culling_band < 0.0f ? 0.0f : culling_band;
    float best = band * band;
    std::uint32_t stack[StackCap];
    int sp = 0;
    float result = 0.0f;
    bool done = false;
    std::uint64_t reference = root_ref;
    while (!done) {
        std::uint64_t next = 0;
        if (ref_leaf(reference)) {
            if (!visit_leaf(tris, nt, reference, x, y, z, best))
                break; // distance zero: result stays 0
        } else if (ref_internal(reference) && ref_offset(reference) < node_count) {
            bvh_node4 const &node = nodes[ref_offset(reference)];
            float cd[4];
            std::uint64_t cr[4];
            int m = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                bvh_child const ch = node.children[s];
                if (!ref_valid(ch.reference))
                    continue;
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
            bool zero = false;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (s < m && ref_leaf(cr[s]) && cd[s] < best && !visit_leaf(tris, nt, cr[s], x, y, z, best)) {
                    zero = true;
                    break;
                }
            }
            if (zero)
                break;
#pragma unroll
            for (int s = 3; s >= 0; --s) {
                if (s >= m || !ref_internal(cr[s]) || cd[s] >= best)
                    continue;
                if (next != 0) {
                    if (sp >= StackCap) {
                        result = __builtin_nanf("");
                        done = true;
                        break;
                    }
                    stack[sp++] = static_cast<std::uint32_t>(ref_offset(next));
                }
                next = cr[s];
            }
            if (done)
                break;
        }
        if (next != 0) {
            reference = next;
            continue;
        }
        if (sp == 0) {
            result = sqrtf(best);
            break;
        }
        reference = encode_internal(stack[--sp]);
    }
    out[q] = result;
}

template <int Cap>
void launch_distance(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, std::uint32_t const *order,
    float *out, std::size_t n, float band, hipStream_t stream
) {
    distance_kernel<Cap><<<grid_for(n, k_block), k_block, 0, stream>>>(
        bvh.nodes.ptr, bvh.node_count, bvh.triangles.ptr, bvh.triangle_count, bvh.root.reference, qx, qy, qz,
        order, out, n, band
    );
    GWN_LAUNCH_CHECK();
}

} // namespace

void unsigned_distance(
    gwn_bvh_handle const &bvh, float const *qx, float const *qy, float const *qz, float *out, std::size_t n,
    float culling_band, query_launch const &cfg, hipStream_t stream
) {
    if (n == 0)
        return;
    bool const sort = cfg.sort_queries > 0 || (cfg.sort_queries < 0 && n >= (1u << 16));
    device_buffer<std::uint32_t> order;
    if (sort) {
        order.allocate(n, stream);
        query_order(bvh, qx, qy, qz, order.ptr, n, stream);
    }
    int const need = static_cast<int>(bvh.internal_stack_bound);
    if (need <= 32)
        launch_distance<32>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, stream);
    else if (need <= 64)
        launch_distance<64>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, stream);
    else if (need <= 128)
        launch_distance<128>(bvh, qx, qy, qz, order.ptr, out, n, culling_band, stream);
    else
        throw_invalid("Traversal stack bound exceeds the largest compiled stack (128).");
}

} // namespace gwn
