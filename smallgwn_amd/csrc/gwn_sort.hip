// gwn_sort.hip - stable LSD radix sort of (key, uint32 value) pairs for gfx950.
//
// Replaces the CUB DeviceRadixSort::SortPairs calls of the reference build
// (include/gwn/detail/gwn_bvh_build_common.cuh:325-347, gwn_bvh_reorder.cuh:104-116).
//
// Design (reduce-then-scan, 8-bit digits, 4096-key tiles of one 256-thread block):
//   1. one histogram sweep counts every digit of every pass at once, and one scan turns the
//      counts into per-pass digit bases;
//   2. per pass, a count kernel records each tile's digit counts (digit-major table), a scan
//      kernel (one block per digit) turns each digit's row into global tile offsets, and a
//      scatter kernel ranks the tile with wave-wide ballot "match" masks (stable: rank order =
//      input order), stages it in LDS in digit order and writes equal-digit runs contiguously.
// No workgroup waits on another: every cross-tile dependency is a kernel boundary.  The table
// costs 1 KiB per tile and one extra read of the keys per pass; a single-pass look-back design
// serializes on one status round trip per tile, which on a multi-XCD part costs more than the
// whole pass's traffic.
#include "gwn_core.hpp"

namespace gwn {
namespace {

constexpr int k_sort_block = 256;
constexpr int k_sort_waves = k_sort_block / 64;
constexpr int k_radix = 256;
constexpr int k_items = 16;
constexpr int k_tile = k_sort_block * k_items;
constexpr int k_max_passes = 8;

constexpr std::size_t k_max_sort_count = (std::size_t(1) << 31) - 1u;

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

// Stable in-wave ranking of the tile's items.  Item `it` of lane l is element
// wave*64*k_items + it*64 + l of the tile, so ranking items in (it, lane) order ranks elements
// in input order.  For each item, `match` is the set of lanes holding the same digit; the lowest
// such lane advances the wave's digit counter.
template <class Key>
__device__ inline void rank_tile_items(
    Key const (&key)[k_items], int tile_count, int wave_base, int shift, unsigned digit_mask,
    std::uint32_t *wave_count, std::uint32_t (&rank)[k_items]
) {
    int const lane = threadIdx.x & 63;
    std::uint64_t const lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
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
            std::uint32_t const old = wave_count[d];
            r = old + static_cast<std::uint32_t>(__popcll(match & lt_mask));
            int const leader = __ffsll(static_cast<long long>(match)) - 1;
            if (lane == leader)
                wave_count[d] = old + static_cast<std::uint32_t>(__popcll(match));
        }
        rank[it] = r;
    }
}

// counts[d * tiles + t] = number of keys of tile t whose digit is d.
template <class Key>
__global__ __launch_bounds__(k_sort_block) void tile_count_kernel(
    Key const *__restrict__ keys, std::size_t n, int shift, unsigned digit_mask, std::uint32_t tiles,
    std::uint32_t *__restrict__ counts
) {
    __shared__ std::uint32_t s_count[k_sort_waves][k_radix];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < k_sort_waves * k_radix; i += k_sort_block)
        (&s_count[0][0])[i] = 0u;
    __syncthreads();
    std::size_t const tile_begin = static_cast<std::size_t>(blockIdx.x) * k_tile;
    std::size_t const remaining = n - tile_begin;
    int const tile_count = remaining < static_cast<std::size_t>(k_tile) ? static_cast<int>(remaining) : k_tile;
    int const wave_base = wave * 64 * k_items;
    Key key[k_items];
#pragma unroll
    for (int it = 0; it < k_items; ++it) {
        int const local = wave_base + it * 64 + lane;
        key[it] = local < tile_count ? keys[tile_begin + local] : Key(0);
    }
    std::uint32_t rank[k_items];
    rank_tile_items<Key>(key, tile_count, wave_base, shift, digit_mask, s_count[wave], rank);
    __syncthreads();
    std::uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < k_sort_waves; ++w)
        total += s_count[w][threadIdx.x];
    counts[static_cast<std::size_t>(threadIdx.x) * tiles + blockIdx.x] = total;
}

// One block per digit: row d of the count table becomes the global output offset of each
// tile's run of digit d (digit base + exclusive prefix over earlier tiles).
__global__ __launch_bounds__(k_sort_block) void tile_scan_kernel(
    std::uint32_t *__restrict__ counts, std::uint32_t tiles, std::uint32_t const *__restrict__ digit_base
) {
    __shared__ std::uint32_t s_tot[k_sort_waves];
    std::uint32_t *row = counts + static_cast<std::size_t>(blockIdx.x) * tiles;
    std::uint32_t carry = digit_base[blockIdx.x];
    for (std::uint32_t c = 0; c < tiles; c += k_sort_block) {
        std::uint32_t const i = c + threadIdx.x;
        std::uint32_t const v = i < tiles ? row[i] : 0u;
        std::uint32_t total;
        std::uint32_t const e = block_exclusive_scan(v, s_tot, total);
        if (i < tiles)
            row[i] = carry + e;
        carry += total;
    }
}

template <class Key> struct scatter_smem {
    Key keys[k_tile];
    std::uint32_t vals[k_tile];
    std::uint32_t wave_count[k_sort_waves][k_radix];
    std::uint32_t tile_start[k_radix];
    std::uint32_t digit_base[k_radix];
    std::uint32_t wave_totals[k_sort_waves];
};

template <class Key>
__global__ __launch_bounds__(k_sort_block) void scatter_kernel(
    Key const *__restrict__ keys_in, Key *__restrict__ keys_out,
    std::uint32_t const *__restrict__ vals_in, std::uint32_t *__restrict__ vals_out, std::size_t n,
    int shift, unsigned digit_mask, std::uint32_t tiles, std::uint32_t const *__restrict__ offsets
) {
    __shared__ scatter_smem<Key> s;
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < k_sort_waves * k_radix; i += k_sort_block)
        (&s.wave_count[0][0])[i] = 0u;
    __syncthreads();
    std::uint32_t const tile = blockIdx.x;
    std::size_t const tile_begin = static_cast<std::size_t>(tile) * k_tile;
    std::size_t const remaining = n - tile_begin;
    int const tile_count = remaining < static_cast<std::size_t>(k_tile) ? static_cast<int>(remaining) : k_tile;

    Key key[k_items];
    std::uint32_t val[k_items];
    std::uint32_t rank[k_items];
    int const wave_base = wave * 64 * k_items;
#pragma unroll
    for (int it = 0; it < k_items; ++it) {
        int const local = wave_base + it * 64 + lane;
        bool const valid = local < tile_count;
        key[it] = valid ? keys_in[tile_begin + local] : Key(0);
        val[it] = valid ? vals_in[tile_begin + local] : 0u;
    }
    rank_tile_items<Key>(key, tile_count, wave_base, shift, digit_mask, s.wave_count[wave], rank);
    __syncthreads();

    // One thread per digit: per-wave prefixes and the tile's digit total.
    unsigned const digit = threadIdx.x;
    std::uint32_t tile_total = 0;
#pragma unroll
    for (int w = 0; w < k_sort_waves; ++w) {
        std::uint32_t const c = s.wave_count[w][digit];
        s.wave_count[w][digit] = tile_total;
        tile_total += c;
    }
    std::uint32_t block_total;
    std::uint32_t const tile_start = block_exclusive_scan(tile_total, s.wave_totals, block_total);
    s.tile_start[digit] = tile_start;
    s.digit_base[digit] = offsets[static_cast<std::size_t>(digit) * tiles + tile] - tile_start;
    __syncthreads();

    // Stage the tile in digit order, then write each digit run contiguously.
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

// Scratch of one sort: digit bases [passes][256], then the [256][tiles] tile table reused by
// every pass.
std::size_t sort_base_words(int begin_bit, int end_bit) {
    return static_cast<std::size_t>((end_bit - begin_bit + 7) / 8) * k_radix;
}

std::size_t sort_table_words(std::size_t n) { return (n + k_tile - 1) / k_tile * k_radix; }

template <class Key>
void radix_sort_pairs(
    Key *keys, Key *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt, std::size_t n,
    int begin_bit, int end_bit, std::uint32_t *workspace, Key **keys_out, std::uint32_t **vals_out,
    hipStream_t stream
) {
    *keys_out = keys;
    *vals_out = vals;
    if (n <= 1 || end_bit <= begin_bit)
        return;
    if (n > k_max_sort_count)
        throw_invalid("radix sort: element count exceeds 2^31 - 1.");
    int const passes = (end_bit - begin_bit + 7) / 8;
    if (passes > k_max_passes)
        throw_invalid("radix sort: bit range exceeds the key width.");
    std::size_t const tiles = (n + k_tile - 1) / k_tile;
    std::size_t const base_words = sort_base_words(begin_bit, end_bit);
    device_buffer<std::uint32_t> own;
    if (workspace == nullptr) {
        own.allocate(base_words + sort_table_words(n), stream);
        workspace = own.ptr;
    }
    // The digit bases accumulate atomically: clear them on every call (a fill kernel).
    fill_u32(workspace, 0u, base_words, stream);
    std::uint32_t *hist = workspace;
    std::uint32_t *table = workspace + base_words;

    unsigned const hist_grid = static_cast<unsigned>(tiles < 2048 ? tiles : 2048);
    histogram_kernel<Key><<<hist_grid, k_sort_block, 0, stream>>>(keys, n, begin_bit, end_bit, passes, hist);
    GWN_LAUNCH_CHECK();
    scan_histogram_kernel<<<1, k_sort_block, 0, stream>>>(hist, passes);
    GWN_LAUNCH_CHECK();

    Key *kin = keys, *kout = keys_alt;
    std::uint32_t *vin = vals, *vout = vals_alt;
    auto const ntiles = static_cast<std::uint32_t>(tiles);
    for (int p = 0; p < passes; ++p) {
        int const shift = begin_bit + 8 * p;
        int const width = end_bit - shift < 8 ? end_bit - shift : 8;
        unsigned const mask = (1u << width) - 1u;
        tile_count_kernel<Key><<<ntiles, k_sort_block, 0, stream>>>(kin, n, shift, mask, ntiles, table);
        GWN_LAUNCH_CHECK();
        tile_scan_kernel<<<k_radix, k_sort_block, 0, stream>>>(table, ntiles, hist + p * k_radix);
        GWN_LAUNCH_CHECK();
        scatter_kernel<Key><<<ntiles, k_sort_block, 0, stream>>>(kin, kout, vin, vout, n, shift, mask, ntiles, table);
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
    radix_sort_pairs<std::uint64_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, nullptr, keys_out, vals_out,
                                    stream);
}

void radix_sort_pairs_u64_ws(
    std::uint64_t *keys, std::uint64_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t *workspace, std::uint64_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
) {
    if (workspace == nullptr)
        throw_internal("radix sort: null workspace.");
    radix_sort_pairs<std::uint64_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, workspace, keys_out,
                                    vals_out, stream);
}

void radix_sort_pairs_u32(
    std::uint32_t *keys, std::uint32_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t **keys_out, std::uint32_t **vals_out,
    hipStream_t stream
) {
    radix_sort_pairs<std::uint32_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, nullptr, keys_out, vals_out,
                                    stream);
}

std::size_t radix_sort_workspace_words(std::size_t n, int begin_bit, int end_bit) {
    return sort_base_words(begin_bit, end_bit) + sort_table_words(n);
}

void radix_sort_pairs_u32_ws(
    std::uint32_t *keys, std::uint32_t *keys_alt, std::uint32_t *vals, std::uint32_t *vals_alt,
    std::size_t n, int begin_bit, int end_bit, std::uint32_t *workspace, std::uint32_t **keys_out,
    std::uint32_t **vals_out, hipStream_t stream
) {
    if (workspace == nullptr)
        throw_internal("radix sort: null workspace.");
    radix_sort_pairs<std::uint32_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, workspace, keys_out,
                                    vals_out, stream);
}

} // namespace gwn
