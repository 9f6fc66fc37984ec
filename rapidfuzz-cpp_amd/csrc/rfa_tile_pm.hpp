// rfa_tile_pm.hpp -- the query tile's pattern-match table: PM[symbol row][lane]
// for up to 64 queries of at most 64 code units, one query per lane
// (PatternMatchVector.hpp:62-134 semantics, symbols compacted to dense rows).
// Shared by the tile kernels of the distance metrics and of partial_ratio.
#pragma once
#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

template <typename CharT>
struct SymLookup;

// 8-bit units: direct 256-entry row table.
template <>
struct SymLookup<uint8_t> {
    const uint8_t* rowmap;
    __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return rowmap[c]; }
};

template <>
struct SymLookup<uint16_t> {
    HashMap map;
    __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return map.find(c); }
};

template <>
struct SymLookup<uint32_t> {
    HashMap map;
    __device__ __forceinline__ uint32_t operator()(uint32_t c) const { return map.find(c); }
};

// Threads of a tile workgroup: 8 waves share one tile's PM table, so a CU
// holds twice the waves a 4-wave workgroup would at the same LDS footprint.
constexpr uint32_t TILE_WAVES = 8;
constexpr uint32_t TILE_THREADS = TILE_WAVES * 64;

// LDS after the PM rows: the symbol map (row table or hash) and the scan scratch.
// A multiple of 8 bytes for every hash size.
__host__ __device__ constexpr uint32_t tile_extra_bytes(int char_width, uint32_t hash_bits)
{
    return char_width == 1 ? 256 * 4 + 256 + 4 * TILE_WAVES : (1u << hash_bits) * 10 + 16 + 4 * TILE_WAVES;
}

// LDS after the symbol map: two 64-bit words per lane of length-settled flags.
constexpr uint32_t TILE_SETTLE_BYTES = 64 * 2 * 8;

// Builds the tile's PM rows in LDS.  Returns the lookup functor.
template <typename CharT, typename W>
__device__ __forceinline__ SymLookup<CharT> build_tile_pm(const CharT* __restrict__ q, uint64_t qo, uint32_t m,
                                                          W* pm, uint8_t* extra, const TileArgs& a)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    SymLookup<CharT> look;
    uint32_t R;
    if constexpr (sizeof(CharT) == 1) {
        uint32_t* flags = (uint32_t*)extra;         // 256
        uint8_t* rowmap = (uint8_t*)(flags + 256);  // 256
        uint32_t* tmp = (uint32_t*)(rowmap + 256);  // TILE_WAVES
        if (tid < 256) flags[tid] = 0;
        __syncthreads();
        for (uint32_t j = wave; j < m; j += TILE_WAVES) flags[q[qo + j]] = 1;
        __syncthreads();
        uint32_t f = tid < 256 ? flags[tid] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan<TILE_WAVES>(f, tmp, &total);
        if (tid < 256) rowmap[tid] = f ? (uint8_t)(ex + 1) : 0;
        R = total + 1;
        look.rowmap = rowmap;
    }
    else {
        const uint32_t H = 1u << a.hash_bits;
        uint64_t* keys = (uint64_t*)extra;
        uint16_t* rows = (uint16_t*)(keys + H);
        uint32_t* tmp = (uint32_t*)(rows + H + 8);
        for (uint32_t i = tid; i < H; i += TILE_THREADS) keys[i] = EMPTY_KEY;
        __syncthreads();
        HashMap map{keys, rows, a.hash_bits};
        for (uint32_t j = wave; j < m; j += TILE_WAVES) map.insert((uint32_t)q[qo + j]);
        __syncthreads();
        uint32_t base = 0;
        for (uint32_t s0 = 0; s0 < H; s0 += TILE_THREADS) {
            uint32_t i = s0 + tid;
            uint32_t f = (i < H && keys[i] != EMPTY_KEY) ? 1u : 0u;
            uint32_t total;
            uint32_t ex = block_exclusive_scan<TILE_WAVES>(f, tmp, &total);
            if (f) rows[i] = (uint16_t)(base + ex + 1);
            base += total;
        }
        R = base + 1;
        look.map = map;
    }
    __syncthreads();
    for (uint32_t i = tid; i < R * 64; i += TILE_THREADS) pm[i] = 0;
    __syncthreads();
    for (uint32_t j = wave; j < m; j += TILE_WAVES) {
        uint32_t r = look((uint32_t)q[qo + j]);
        if constexpr (sizeof(W) == 8)
            atomicOr((unsigned long long*)&pm[r * 64 + lane], 1ull << j);
        else
            atomicOr((unsigned int*)&pm[r * 64 + lane], 1u << j);
    }
    __syncthreads();
    return look;
}

} // namespace rfa
