// gwn_lookback.hpp - single-pass decoupled look-back prefix scan across workgroups (gfx950).
//
// A status word holds a 2-bit flag in bits 30-31 (0 = not yet published, 1 = tile aggregate,
// 2 = inclusive prefix) and a 30-bit value.  The value travels inside the word, so relaxed
// agent-scope atomics (sc1 loads and stores, coherent across the XCD L2s) suffice and no L2
// write-back or invalidate is paid per step.  Tiles must be numbered in the order their
// workgroups started (a claimed tile counter), so every predecessor is resident or finished and
// each wait ends; the spin cap only turns a broken invariant into an error instead of a hang.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gwn {

constexpr std::uint32_t k_lb_aggregate = 1u << 30;
constexpr std::uint32_t k_lb_inclusive = 2u << 30;
constexpr std::uint32_t k_lb_value_mask = (1u << 30) - 1u;
constexpr std::uint32_t k_lb_spin_limit = 1u << 22;

__device__ inline std::uint32_t lb_load(std::uint32_t const *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_store(std::uint32_t *p, std::uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline std::uint32_t wave_sum_u32(std::uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive prefix of `tile` over all earlier tiles.  Called by all 64 lanes of one wave with the
// same arguments; returns the prefix to every lane and publishes the tile's inclusive value.
// Each step reads the 64 nearest unresolved predecessors at once (lane j reads tile end-1-j).
__device__ inline std::uint32_t wave_lookback(
    std::uint32_t *status, std::uint32_t tile, std::uint32_t tile_total, std::uint32_t *error
) {
    int const lane = threadIdx.x & 63;
    if (lane == 0)
        lb_store(status + tile, (tile == 0 ? k_lb_inclusive : k_lb_aggregate) | tile_total);
    if (tile == 0)
        return 0u;
    std::uint32_t exclusive = 0;
    std::int64_t end = tile;
    std::uint32_t spins = 0;
    for (;;) {
        std::int64_t const idx = end - 1 - lane;
        std::uint32_t const st = idx >= 0 ? lb_load(status + idx) : k_lb_inclusive;
        std::uint32_t const flag = st >> 30;
        std::uint64_t const incl = __ballot(flag == 2u);
        std::uint64_t const pending = __ballot(flag == 0u);
        // Lanes up to and including the nearest inclusive predecessor (all lanes if none).
        std::uint64_t const upto = incl ? ((incl & (~incl + 1ull)) << 1) - 1ull : ~0ull;
        if (pending & upto) {
            if (++spins > k_lb_spin_limit) {
                if (lane == 0)
                    atomicOr(error, 4u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        exclusive += wave_sum_u32(((upto >> lane) & 1ull) ? (st & k_lb_value_mask) : 0u);
        if (incl)
            break;
        end -= 64;
    }
    if (lane == 0)
        lb_store(status + tile, k_lb_inclusive | (exclusive + tile_total));
    return exclusive;
}

} // namespace gwn
