// rfa_plan.hip -- device-side planning of the bucketed pairs schedule.
//
//   classify : per pair, the kernel, the pattern side and a 32-bit sort key
//              (descending order):
//                segmented : Levenshtein / Indel / LCSseq pairs whose shorter
//                            string has <= 4096 units; the pattern is the
//                            shorter string, or the longer one when it has
//                            512-4096 units; bit 28 for texts of >= 512 units
//                            (long sequential chains, scheduled first), class
//                            c = log2(lanes per pair) in bits 25-27,
//                            Levenshtein flag in bit 24, text length below
//                OSA short : 2^30 (+ 2^29 for 33-64 units) + text length
//                block     : the rest; key 2^31 + words * text length (cost,
//                            saturated)
//   sort     : one descending radix sort of (key, pair|swap<<31) -- block pairs
//              come first, most expensive first (longest-processing-time order
//              for the block kernel's queue); the segmented classes follow from
//              widest to narrowest, grouped by metric and text length so the
//              lanes of a wave finish together.
//   split    : one thread finds the range boundaries and resets the block queue.
// The host never reads the plan back, so a call costs O(1) host work.
#include <hipcub/hipcub.hpp>

#include "rfa_internal.hpp"

namespace rfa {

constexpr uint32_t BLOCK_FLAG = 1u << 31;
constexpr uint32_t OSA_FLAG = 1u << 30;
constexpr uint32_t S64_FLAG = 1u << 29;
constexpr uint32_t CHAIN_FLAG = 1u << 28; // segmented pairs with long texts
constexpr uint32_t LEV_FLAG = 1u << 24;
constexpr uint32_t SEG_MAX_UNITS = 4096;

__device__ __forceinline__ uint32_t pow2_at_least(uint64_t x)
{
    uint32_t c = 0;
    while ((1ull << c) < x) ++c;
    return c;
}

__device__ __forceinline__ uint32_t sat(uint64_t v, uint32_t bits) { return (uint32_t)(v < (1ull << bits) ? v : (1ull << bits) - 1); }

__global__ void k_classify(const uint64_t* __restrict__ aoff, const uint64_t* __restrict__ boff,
                           const int32_t* __restrict__ metric, uint64_t n, uint32_t* __restrict__ keys,
                           uint32_t* __restrict__ vals)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t la = aoff[i + 1] - aoff[i], lb = boff[i + 1] - boff[i];
        bool pat_b = lb < la; // the pattern is the shorter string ...
        const uint64_t s = pat_b ? lb : la, l = pat_b ? la : lb;
        const int32_t met = metric[i];
        const uint32_t lev = met == RFA_LEVENSHTEIN ? LEV_FLAG : 0u;
        uint32_t key;
        if (met != RFA_OSA && l <= SEG_MAX_UNITS && l >= SEG_CHAIN_STEPS) {
            // ... except in the segmented kernel when the longer string has >= 512 units:
            // then the longer string is the pattern, spread over up to 64 lanes, and the
            // sequential chain is the shorter string (a 60-unit string against a
            // 3000-unit one: 64 lanes x 60 steps instead of 1 lane x 3000 steps)
            pat_b = !pat_b;
            key = (s >= SEG_CHAIN_STEPS ? CHAIN_FLAG : 0u) | (pow2_at_least((l + 63) / 64) << 25) | lev | (uint32_t)s;
        }
        else if (met != RFA_OSA && s <= SEG_MAX_UNITS) {
            key = (l >= SEG_CHAIN_STEPS ? CHAIN_FLAG : 0u) | (pow2_at_least((s + 63) / 64) << 25) | lev | sat(l, 24);
        }
        else if (met == RFA_OSA && s <= 64) {
            key = OSA_FLAG | (s > 32 ? S64_FLAG : 0u) | sat(l, 29);
        }
        else {
            key = BLOCK_FLAG | sat(((s + 63) / 64) * l, 31);
        }
        keys[i] = key;
        vals[i] = (uint32_t)i | (pat_b ? 0x80000000u : 0u);
    }
}

// First index in [lo, hi) whose key is below thr (keys sorted descending).
__device__ uint64_t first_below(const uint32_t* keys, uint64_t lo, uint64_t hi, uint32_t thr)
{
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (keys[mid] >= thr)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_split(const uint32_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ counters)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t e = first_below(keys, 0, n, BLOCK_FLAG);
    counters[1] = (uint32_t)e;
    e = first_below(keys, e, n, OSA_FLAG | S64_FLAG);
    counters[2] = (uint32_t)e;
    e = first_below(keys, e, n, OSA_FLAG);
    counters[3] = (uint32_t)e;
    for (uint32_t r = 0; r < 2 * SEG_CLASSES; ++r) {
        const uint32_t c = SEG_CLASSES - 1 - r % SEG_CLASSES;
        e = first_below(keys, e, n, (r < SEG_CLASSES ? CHAIN_FLAG : 0u) | (c << 25));
        counters[4 + r] = (uint32_t)e;
    }
    counters[0] = 0;
}

size_t plan_temp_bytes(uint64_t n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
    return bytes;
}

hipError_t launch_plan(const uint64_t* aoff, const uint64_t* boff, const int32_t* metric, uint64_t n,
                       uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_in, uint32_t* vals_out, void* temp,
                       size_t temp_bytes, uint32_t* counters, hipStream_t s)
{
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(256), 0, s, aoff, boff, metric, n, keys_in, vals_in);
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                                (int)n, 0, 32, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_split, dim3(1), dim3(64), 0, s, keys_out, n, counters);
    return hipGetLastError();
}

} // namespace rfa
