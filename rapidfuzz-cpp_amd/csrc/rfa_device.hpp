// rfa_device.hpp -- wave-level helpers and the per-tile symbol maps (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rfa {

constexpr int WAVE = 64;
constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t readlane(uint32_t v, uint32_t l) { return __builtin_amdgcn_readlane(v, l); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t blsi(uint64_t x) { return x & (0ull - x); }

// Multiplicative hash of a 32-bit symbol into a table of 2^bits slots.
__device__ __forceinline__ uint32_t sym_hash(uint32_t c, uint32_t bits) { return (c * 0x9E3779B1u) >> (32 - bits); }

// Open-addressing symbol -> row map in LDS for 16/32-bit code units.
// Rows are dense ids 1..R-1 (row 0 is the all-zero PM row of absent symbols).
struct HashMap {
    uint64_t* keys; // [1 << bits], EMPTY_KEY when free
    uint16_t* rows; // [1 << bits]
    uint32_t bits;

    __device__ __forceinline__ uint32_t find(uint32_t c) const
    {
        uint32_t mask = (1u << bits) - 1;
        uint32_t h = sym_hash(c, bits);
        for (;;) {
            uint64_t k = keys[h];
            if (k == (uint64_t)c) return rows[h];
            if (k == EMPTY_KEY) return 0;
            h = (h + 1) & mask;
        }
    }
    // Insert; returns the slot.  Concurrent inserts of one key resolve to one slot.
    __device__ __forceinline__ uint32_t insert(uint32_t c)
    {
        uint32_t mask = (1u << bits) - 1;
        uint32_t h = sym_hash(c, bits);
        for (;;) {
            uint64_t prev = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)EMPTY_KEY,
                                      (unsigned long long)c);
            if (prev == EMPTY_KEY || prev == (uint64_t)c) return h;
            h = (h + 1) & mask;
        }
    }
};

// K consecutive code units from p (any alignment: the loads are dword loads of
// the covering bytes).  Callers that read past a string's end rely on the
// 64-byte tail padding of every string set.
template <typename CharT, uint32_t K>
__device__ __forceinline__ void load_units(const CharT* p, uint32_t* out)
{
    CharT tmp[K];
    __builtin_memcpy(tmp, p, sizeof(tmp));
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) out[k] = (uint32_t)tmp[k];
}

// Exclusive prefix sum of one value per thread over a block of NW waves.
// `tmp` holds at least NW uint32 of LDS.  Returns the exclusive sum; *total gets the block sum.
template <uint32_t NW>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* tmp, uint32_t* total)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (uint32_t w = 0; w < NW; ++w) {
        uint32_t t = tmp[w];
        if (w < wave) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

} // namespace rfa
