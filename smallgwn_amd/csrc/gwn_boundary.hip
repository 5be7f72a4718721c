// gwn_boundary.hip - algebraic boundary chain of an oriented triangle mesh on gfx950.
//
// Reference: include/gwn/detail/gwn_boundary_impl.cuh:21-354, include/gwn/gwn_boundary.cuh.
//
// Every oriented triangle edge (a -> b) becomes the undirected key (lo, hi) packed as
// lo << 32 | hi with orientation +1 when a < b, -1 when b < a and 0 for a degenerate edge.  One
// stable radix sort of the 64-bit keys orders them by (lo, hi) -- the order the reference reaches
// with two 32-bit passes.  The first row of each key group sums the group's orientations; a
// nonzero sum is one directed boundary row (lo -> hi for a positive sum, hi -> lo otherwise) with
// multiplicity |sum|.  A look-back scan of the group flags places the rows, in key order.
#include "gwn_core.hpp"
#include "gwn_lookback.hpp"

namespace gwn {
namespace {

constexpr int k_block = 256;

__global__ __launch_bounds__(k_block) void boundary_edges_kernel(
    std::uint32_t const *__restrict__ i0, std::uint32_t const *__restrict__ i1,
    std::uint32_t const *__restrict__ i2, std::size_t nt, std::uint64_t *__restrict__ keys,
    std::uint32_t *__restrict__ orientation
) {
    std::size_t const t = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (t >= nt)
        return;
    std::uint32_t const v[3] = {i0[t], i1[t], i2[t]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        std::uint32_t const a = v[k], b = v[(k + 1) % 3];
        std::uint32_t const lo = a < b ? a : b, hi = a < b ? b : a;
        keys[3 * t + k] = (static_cast<std::uint64_t>(lo) << 32) | hi;
        orientation[3 * t + k] = static_cast<std::uint32_t>(a < b ? 1 : (b < a ? -1 : 0));
    }
}

// flag[e] = 1 when row e starts a key group whose orientation sum is nonzero; that sum goes to
// group_sum[e].
__global__ __launch_bounds__(k_block) void boundary_heads_kernel(
    std::uint64_t const *__restrict__ keys, std::uint32_t const *__restrict__ orientation, std::size_t m,
    std::uint32_t *__restrict__ flag, std::int32_t *__restrict__ group_sum
) {
    std::size_t const e = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (e >= m)
        return;
    if (e > 0 && keys[e] == keys[e - 1]) {
        flag[e] = 0u;
        return;
    }
    std::int32_t sum = static_cast<std::int32_t>(orientation[e]);
    for (std::size_t j = e + 1; j < m && keys[j] == keys[e]; ++j)
        sum += static_cast<std::int32_t>(orientation[j]);
    flag[e] = sum != 0 ? 1u : 0u;
    group_sum[e] = sum;
}

// Exclusive scan of the flags over 256-row tiles with a wave-parallel look-back; the last tile
// also writes the total.
__global__ __launch_bounds__(k_block) void boundary_scan_kernel(
    std::uint32_t const *__restrict__ flag, std::size_t m, std::uint32_t *__restrict__ offset,
    std::uint32_t *status, std::uint32_t *tile_counter, std::uint32_t *total, std::uint32_t *error
) {
    __shared__ std::uint32_t s_tile, s_exclusive;
    __shared__ std::uint32_t s_wave[k_block / 64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    if (threadIdx.x == 0)
        s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    std::uint32_t const tile = s_tile;
    std::size_t const tiles = (m + k_block - 1) / k_block;
    if (tile >= tiles)
        return;
    std::size_t const e = static_cast<std::size_t>(tile) * k_block + threadIdx.x;
    std::uint32_t const f = e < m ? flag[e] : 0u;
    std::uint32_t incl = f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        std::uint32_t const v = __shfl_up(incl, o, 64);
        if (lane >= o)
            incl += v;
    }
    if (lane == 63)
        s_wave[wave] = incl;
    __syncthreads();
    std::uint32_t wave_prefix = 0, block_total = 0;
    for (int w = 0; w < k_block / 64; ++w) {
        if (w < wave)
            wave_prefix += s_wave[w];
        block_total += s_wave[w];
    }
    if (wave == 0) {
        std::uint32_t const ex = wave_lookback(status, tile, block_total, error);
        if (lane == 0)
            s_exclusive = ex;
    }
    __syncthreads();
    if (e < m)
        offset[e] = s_exclusive + wave_prefix + incl - f;
    if (threadIdx.x == 0 && tile == tiles - 1)
        *total = s_exclusive + block_total;
}

__global__ __launch_bounds__(k_block) void boundary_emit_kernel(
    std::uint64_t const *__restrict__ keys, std::size_t m, std::uint32_t const *__restrict__ flag,
    std::int32_t const *__restrict__ group_sum, std::uint32_t const *__restrict__ offset,
    std::uint32_t *__restrict__ start, std::uint32_t *__restrict__ end, std::uint64_t *__restrict__ multiplicity
) {
    std::size_t const e = static_cast<std::size_t>(blockIdx.x) * k_block + threadIdx.x;
    if (e >= m || flag[e] == 0u)
        return;
    std::uint32_t const lo = static_cast<std::uint32_t>(keys[e] >> 32);
    std::uint32_t const hi = static_cast<std::uint32_t>(keys[e]);
    std::int32_t const sum = group_sum[e];
    std::uint32_t const o = offset[e];
    start[o] = sum > 0 ? lo : hi;
    end[o] = sum > 0 ? hi : lo;
    multiplicity[o] = static_cast<std::uint64_t>(sum > 0 ? sum : -sum);
}

int bits_for_value(std::uint64_t v) {
    int b = 0;
    while (b < 64 && (std::uint64_t(1) << b) <= v)
        ++b;
    return b;
}

} // namespace

void build_boundary_chain(gwn_geometry_handle const &g, gwn_boundary_handle &out, hipStream_t stream) {
    out.mesh_vertex_count = g.vertex_count;
    out.mesh_triangle_count = g.triangle_count;
    out.edge_count = 0;
    out.built = true;
    std::size_t const m = 3 * g.triangle_count;
    if (m == 0)
        return;
    device_buffer<std::uint64_t> keys(m, stream), keys_alt(m, stream);
    device_buffer<std::uint32_t> orient(m, stream), orient_alt(m, stream);
    boundary_edges_kernel<<<grid_for(g.triangle_count, k_block), k_block, 0, stream>>>(
        g.i0.ptr, g.i1.ptr, g.i2.ptr, g.triangle_count, keys.ptr, orient.ptr
    );
    GWN_LAUNCH_CHECK();
    std::uint64_t *sk = nullptr;
    std::uint32_t *so = nullptr;
    int const end_bit = 32 + bits_for_value(g.vertex_count > 0 ? g.vertex_count - 1 : 0);
    radix_sort_pairs_u64(keys.ptr, keys_alt.ptr, orient.ptr, orient_alt.ptr, m, 0, end_bit, &sk, &so, stream);

    std::size_t const tiles = (m + k_block - 1) / k_block;
    device_buffer<std::uint32_t> flag(m, stream), offset(m, stream);
    device_buffer<std::int32_t> group_sum(m, stream);
    // Scratch words: tile statuses, tile counter, total, error.
    device_buffer<std::uint32_t> words(tiles + 3, stream);
    fill_u32(words.ptr, 0u, tiles + 3, stream);
    boundary_heads_kernel<<<grid_for(m, k_block), k_block, 0, stream>>>(sk, so, m, flag.ptr, group_sum.ptr);
    GWN_LAUNCH_CHECK();
    std::uint32_t *status = words.ptr;
    std::uint32_t *counter = words.ptr + tiles;
    std::uint32_t *total = counter + 1;
    std::uint32_t *error = counter + 2;
    boundary_scan_kernel<<<static_cast<unsigned>(tiles), k_block, 0, stream>>>(
        flag.ptr, m, offset.ptr, status, counter, total, error
    );
    GWN_LAUNCH_CHECK();
    std::uint32_t host[2] = {0, 0};
    GWN_HIP(hipMemcpyAsync(host, total, 2 * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream));
    GWN_HIP(hipStreamSynchronize(stream));
    if (host[1] != 0u || host[0] > m)
        throw_internal("Boundary chain scan did not complete.");
    std::size_t const rows = host[0];
    out.start.allocate(rows, stream);
    out.end.allocate(rows, stream);
    out.multiplicity.allocate(rows, stream);
    out.edge_count = rows;
    if (rows == 0)
        return;
    boundary_emit_kernel<<<grid_for(m, k_block), k_block, 0, stream>>>(
        sk, m, flag.ptr, group_sum.ptr, offset.ptr, out.start.ptr, out.end.ptr, out.multiplicity.ptr
    );
    GWN_LAUNCH_CHECK();
}

} // namespace gwn
