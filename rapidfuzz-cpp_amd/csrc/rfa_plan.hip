// rfa_plan.hip -- device-side planning of the bucketed pairs schedule.
//
//   classify : per pair, the pattern (the shorter string), its word count and
//              a sort key: long patterns (> 64 units) get 2^40 + words * text
//              length (their cost), short ones get the text length.
//   sort     : one descending radix sort of (key, pair|swap<<31) -- long pairs
//              come first, most expensive first (longest-processing-time
//              order for the block kernel's queue); short pairs follow, grouped
//              by text length so the lanes of a wave finish together.
//   split    : one thread finds the long/short boundary and resets the block
//              kernel's queue counter.
// The host never reads the plan back, so a call costs O(1) host work.
#include <hipcub/hipcub.hpp>

#include "rfa_internal.hpp"

namespace rfa {

constexpr uint64_t LONG_FLAG = 1ull << 40;
constexpr uint64_t S64_FLAG = 1ull << 39;
constexpr uint64_t LOW37 = (1ull << 37) - 1;

// key bits: [40] long pattern (> 64 units) with the cost (words x text length) below;
// otherwise [39] pattern of 33-64 units, [37:38] metric, [0:36] text length.
__global__ void k_classify(const uint64_t* __restrict__ aoff, const uint64_t* __restrict__ boff,
                           const int32_t* __restrict__ metric, uint64_t n, uint64_t* __restrict__ keys,
                           uint32_t* __restrict__ vals)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t la = aoff[i + 1] - aoff[i], lb = boff[i + 1] - boff[i];
        const bool swap = lb < la;
        const uint64_t m = swap ? lb : la, t = swap ? la : lb;
        uint64_t key;
        if (m > 64) {
            const uint64_t cost = ((m + 63) / 64) * t;
            key = LONG_FLAG | (cost < (LONG_FLAG - 1) ? cost : (LONG_FLAG - 1));
        }
        else {
            key = (m > 32 ? S64_FLAG : 0) | ((uint64_t)(metric[i] & 3) << 37) | (t < LOW37 ? t : LOW37);
        }
        keys[i] = key;
        vals[i] = (uint32_t)i | (swap ? 0x80000000u : 0u);
    }
}

// First index in [lo, hi) whose key lacks `flag` (keys sorted descending, flagged first).
__device__ uint64_t first_without(const uint64_t* keys, uint64_t lo, uint64_t hi, uint64_t flag)
{
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (keys[mid] & flag)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_split(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ counters)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t end_long = first_without(keys, 0, n, LONG_FLAG);
    const uint64_t end_s64 = first_without(keys, end_long, n, S64_FLAG);
    counters[0] = 0;
    counters[1] = (uint32_t)end_long;
    counters[2] = (uint32_t)end_s64;
}

size_t plan_temp_bytes(uint64_t n)
{
    size_t bytes = 0;
    hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                 (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 41);
    return bytes;
}

hipError_t launch_plan(const uint64_t* aoff, const uint64_t* boff, const int32_t* metric, uint64_t n,
                       uint64_t* keys_in, uint64_t* keys_out, uint32_t* vals_in, uint32_t* vals_out, void* temp,
                       size_t temp_bytes, uint32_t* counters, hipStream_t s)
{
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(256), 0, s, aoff, boff, metric, n, keys_in, vals_in);
    hipError_t e = hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                                                (int)n, 0, 41, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_split, dim3(1), dim3(64), 0, s, keys_out, n, counters);
    return hipGetLastError();
}

} // namespace rfa
