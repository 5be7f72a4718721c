// gwn_sort.hip - stable LSD radix sort of (key, uint32 value) pairs for gfx950.
//
// Replaces the CUB DeviceRadixSort::SortPairs calls of the reference build
// (include/gwn/detail/gwn_bvh_build_common.cuh:325-347, gwn_bvh_reorder.cuh:104-116).
//
// Design (onesweep, 8-bit digits):
//   1. one histogram sweep counts every digit of every pass at once;
//   2. one scan turns the histograms into per-pass digit offsets;
//   3. one kernel per pass: a 256-thread block (four wave64s) ranks a 4096-key tile with
//      wave-wide ballot "match" masks (stable: rank order = input order), resolves its global
//      digit offsets with decoupled look-back over the preceding tiles, stages the tile in LDS
//      in digit order and writes it out so that equal-digit runs are contiguous stores.
// Tile IDs come from an atomic counter, so every tile a block waits on belongs to a block that
// has already started: the look-back always terminates.
#include "gwn_core.hpp"

namespace gwn {
namespace {

constexpr int k_sort_block = 256;
constexpr int k_sort_waves = k_sort_block / 64;
constexpr int k_radix = 256;
constexpr int k_items = 16;
constexpr int k_tile = k_sort_block * k_items;
constexpr int k_max_passes = 8;

constexpr std::uint32_t k_flag_aggregate = 1u << 30;
constexpr std::uint32_t k_flag_inclusive = 2u << 30;
constexpr std::uint32_t k_value_mask = (1u << 30) - 1u;

__global__ void fill_u32_kernel(std::uint32_t *p, std::uint32_t v, std::size_t n) {
    std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    std::size_t stride = static_cast<std::size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride)
        p[i] = v;
}

template <class Key>
__global__ __launch_bounds__(k_sort_block) void histogram_kernel(
    Key const *__restrict__ keys, std::size_t n, int begin_bit, int end_bit, int passes,
    std::uint32_t *__restrict__ hist
) {
    __shared__ std::uint32_t s_hist[k_max_passes][k_radix];
    for (int i = threadIdx.x; i < k_max_passes * k_radix; i += k_sort_block)
        (&s_hist[0][0])[i] = 0u;
    __syncthreads();
    std::size_t i = static_cast<std::size_t>(blockIdx.x) * k_sort_block + threadIdx.x;
    std::size_t stride = static_cast<std::size_t>(gridDim.x) * k_sort_block;
    for (; i < n; i += stride) {
        Key const k = keys[i];
        for (int p = 0; p < passes; ++p) {
            int const sh = begin_bit + 8 * p;
            int const width = end_bit - sh < 8 ? end_bit - sh : 8;
            unsigned d = static_cast<unsigned>((k >> sh) & Key((1u << width) - 1u));
            atomicAdd(&s_hist[p][d], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < passes; ++p) {
        std::uint32_t v = s_hist[p][threadIdx.x];
        if (v != 0u)
            atomicAdd(&hist[p * k_radix + threadIdx.x], v);
    }
}

// Exclusive scan of one value per thread across a 256-thread block.
__device__ inline std::uint32_t block_exclusive_scan(
    std::uint32_t v, std::uint32_t *s_wave_totals, std::uint32_t &total
) {
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    std::uint32_t incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        std::uint32_t t = __shfl_up(incl, d);
        if (lane >= d)
            incl += t;
    }
    if (lane == 63)
        s_wave_totals[wave] = incl;
    __syncthreads();
    std::uint32_t wave_base = 0;
    total = 0;
    for (int w = 0; w < k_sort_waves; ++w) {
        std::uint32_t t = s_wave_totals[w];
        if (w < wave)
            wave_base += t;
        total += t;
    }
    __syncthreads();
    return wave_base + incl - v;
}

__global__ __launch_bounds__(k_sort_block) void scan_histogram_kernel(std::uint32_t *hist, int passes) {
    __shared__ std::uint32_t s_tot[k_sort_waves];
    for (int p = 0; p < passes; ++p) {
        std::uint32_t total;
        std::uint32_t v = hist[p * k_radix + threadIdx.x];
        std::uint32_t e = block_exclusive_scan(v, s_tot, total);
        hist[p * k_radix + threadIdx.x] = e;
    }
}

template <class Key> struct onesweep_smem {
    Key keys[k_tile];
    std::uint32_t vals[k_tile];
    std::uint32_t wave_count[k_sort_waves][k_radix];
    std::uint32_t tile_start[k_radix];
    std::uint32_t digit_base[k_radix];
    std::uint32_t wave_totals[k_sort_waves];
    std::uint32_t tile_id;
};

template <class Key>
__global__ __launch_bounds__(k_sort_block) void onesweep_kernel(
    Key const *__restrict__ keys_in, Key *__restrict__ keys_out,
    std::uint32_t const *__restrict__ vals_in, std::uint32_t *__restrict__ vals_out, std::size_t n,
    int shift, unsigned digit_mask, std::uint32_t const *__restrict__ pass_offsets,
    std::uint32_t *tile_status, std::uint32_t *tile_counter
) {
    __shared__ onesweep_smem<Key> s;
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;

    if (threadIdx.x == 0)
        s.tile_id = atomicAdd(tile_counter, 1u);
    for (int i = threadIdx.x; i < k_sort_waves * k_radix; i += k_sort_block)
        (&s.wave_count[0][0])[i] = 0u;
    __syncthreads();
    std::uint32_t const tile = s.tile_id;
    std::size_t const tile_begin = static_cast<std::size_t>(tile) * k_tile;
    if (tile_begin >= n)
        return;
    std::size_t const remaining = n - tile_begin;
    int const tile_count = remaining < static_cast<std::size_t>(k_tile) ? static_cast<int>(remaining) : k_tile;

    // Wave-striped load: item i of lane l is element wave*64*k_items + i*64 + l of the tile, so
    // ranking items in (i, lane) order ranks elements in input order (stability).
    Key key[k_items];
    std::uint32_t val[k_items];
    std::uint32_t rank[k_items];
    int const wave_base = wave * 64 * k_items;
    std::uint64_t const lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int it = 0; it < k_items; ++it) {
        int const local = wave_base + it * 64 + lane;
        bool const valid = local < tile_count;
        key[it] = valid ? keys_in[tile_begin + local] : Key(0);
        val[it] = valid ? vals_in[tile_begin + local] : 0u;
    }
#pragma unroll
    for (int it = 0; it < k_items; ++it) {
        int const local = wave_base + it * 64 + lane;
        bool const valid = local < tile_count;
        unsigned const d = static_cast<unsigned>((key[it] >> shift) & Key(digit_mask));
        std::uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool const bit = (d >> b) & 1u;
            std::uint64_t const vote = __ballot(bit);
            match &= bit ? vote : ~vote;
        }
        std::uint32_t r = 0;
        if (valid) {
            std::uint32_t const old = s.wave_count[wave][d];
            r = old + static_cast<std::uint32_t>(__popcll(match & lt_mask));
            int const leader = __ffsll(static_cast<long long>(match)) - 1;
            if (lane == leader)
                s.wave_count[wave][d] = old + static_cast<std::uint32_t>(__popcll(match));
        }
        rank[it] = r;
    }
    __syncthreads();

    // One thread per digit: wave prefixes, tile aggregate, decoupled look-back.
    unsigned const digit = threadIdx.x;
    std::uint32_t tile_total = 0;
#pragma unroll
    for (int w = 0; w < k_sort_waves; ++w) {
        std::uint32_t const c = s.wave_count[w][digit];
        s.wave_count[w][digit] = tile_total;
        tile_total += c;
    }
    std::uint32_t *my_status = tile_status + static_cast<std::size_t>(tile) * k_radix + digit;
    std::uint32_t exclusive = 0;
    if (tile == 0) {
        __hip_atomic_store(my_status, k_flag_inclusive | tile_total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(my_status, k_flag_aggregate | tile_total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        std::int64_t pred = static_cast<std::int64_t>(tile) - 1;
        std::uint32_t spins = 0;
        while (pred >= 0) {
            std::uint32_t const st = __hip_atomic_load(
                tile_status + static_cast<std::size_t>(pred) * k_radix + digit, __ATOMIC_ACQUIRE,
                __HIP_MEMORY_SCOPE_AGENT
            );
            std::uint32_t const flag = st & ~k_value_mask;
            if (flag == 0u) {
                // Every predecessor tile publishes its aggregate before its own look-back, so
                // this wait ends; the cap only bounds the kernel if that invariant were broken.
                if (++spins > (1u << 24))
                    break;
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            exclusive += st & k_value_mask;
            if (flag == k_flag_inclusive)
                break;
            --pred;
        }
        __hip_atomic_store(my_status, k_flag_inclusive | (exclusive + tile_total), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    std::uint32_t block_total;
    std::uint32_t const tile_start = block_exclusive_scan(tile_total, s.wave_totals, block_total);
    s.tile_start[digit] = tile_start;
    s.digit_base[digit] = pass_offsets[digit] + exclusive - tile_start;
    __syncthreads();

    // Stage the tile in digit order.
#pragma unroll
    for (int it = 0; it < k_items; ++it) {
        int const local = wave_base + it * 64 + lane;
        if (local < tile_count) {
            unsigned const d = static_cast<unsigned>((key[it] >> shift) & Key(digit_mask));
            std::uint32_t const pos = s.tile_start[d] + s.wave_count[wave][d] + rank[it];
            s.keys[pos] = key[it];
            s.vals[pos] = val[it];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < tile_count; j += k_sort_block) {
        Key const k = s.keys[j];
        unsigned const d = static_cast<unsigned>((k >> shift) & Key(digit_mask));
        std::size_t const dst = static_cast<std::size_t>(s.digit_base[d]) + static_cast<std::size_t>(j);
        keys_out[dst] = k;
        vals_out[dst] = s.vals[j];
    }
}

template <class Key>
void radix_sort_pairs(
    Key *keys, Key *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt, std::size_t n,
    int begin_bit, int end_bit, Key **keys_out, std::uint32_t **vals_out, hipStream_t stream
) {
    *keys_out = keys;
    *vals_out = vals;
    if (n <= 1 || end_bit <= begin_bit)
        return;
    if (n > static_cast<std::size_t>(k_value_mask))
        throw_invalid("radix sort: element count exceeds 2^30 - 1.");
    int const passes = (end_bit - begin_bit + 7) / 8;
    if (passes > k_max_passes)
        throw_invalid("radix sort: bit range exceeds the key width.");
    std::size_t const tiles = (n + k_tile - 1) / k_tile;
    // Scratch: histograms [passes][256], then per pass a tile counter and tile status words.
    std::size_t const status_words = tiles * k_radix;
    std::size_t const scratch_words = static_cast<std::size_t>(passes) * k_radix +
                                      static_cast<std::size_t>(passes) * (1 + status_words);
    device_buffer<std::uint32_t> scratch(scratch_words, stream);
    fill_u32(scratch.ptr, 0u, scratch_words, stream);
    std::uint32_t *hist = scratch.ptr;
    std::uint32_t *pass_state = scratch.ptr + static_cast<std::size_t>(passes) * k_radix;

    unsigned const hist_grid = static_cast<unsigned>(tiles < 2048 ? tiles : 2048);
    histogram_kernel<Key><<<hist_grid, k_sort_block, 0, stream>>>(keys, n, begin_bit, end_bit, passes, hist);
    GWN_LAUNCH_CHECK();
    scan_histogram_kernel<<<1, k_sort_block, 0, stream>>>(hist, passes);
    GWN_LAUNCH_CHECK();

    Key *kin = keys, *kout = keys_alt;
    std::uint32_t *vin = vals, *vout = vals_alt;
    for (int p = 0; p < passes; ++p) {
        std::uint32_t *counter = pass_state + static_cast<std::size_t>(p) * (1 + status_words);
        int const shift = begin_bit + 8 * p;
        int const width = end_bit - shift < 8 ? end_bit - shift : 8;
        onesweep_kernel<Key><<<static_cast<unsigned>(tiles), k_sort_block, 0, stream>>>(
            kin, kout, vin, vout, n, shift, (1u << width) - 1u, hist + p * k_radix, counter + 1, counter
        );
        GWN_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys_out = kin;
    *vals_out = vin;
}

} // namespace

void fill_u32(std::uint32_t *ptr, std::uint32_t value, std::size_t n, hipStream_t stream) {
    if (n == 0)
        return;
    unsigned grid = grid_for(n, 256);
    if (grid > 4096)
        grid = 4096;
    fill_u32_kernel<<<grid, 256, 0, stream>>>(ptr, value, n);
    GWN_LAUNCH_CHECK();
}

void radix_sort_pairs_u64(
    std::uint64_t *keys, std::uint64_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint64_t **keys_out, std::uint32_t **vals_out,
    hipStream_t stream
) {
    radix_sort_pairs<std::uint64_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, keys_out, vals_out, stream);
}

void radix_sort_pairs_u32(
    std::uint32_t *keys, std::uint32_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t **keys_out, std::uint32_t **vals_out,
    hipStream_t stream
) {
    radix_sort_pairs<std::uint32_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, keys_out, vals_out, stream);
}

} // namespace gwn
