// rfa_plan.hip -- device-side planning of the bucketed pairs schedule: a
// counting sort of the pairs into bins whose ascending order is the order the
// kernels take them in.
//
//   bins (per pair, from its two lengths and its metric):
//     block     : patterns (the shorter string) of > 4096 units, and OSA
//                 patterns of > 64 units; 256 bins of words x text length on a
//                 log scale (4 bins per power of two), most expensive first
//                 (longest-processing-time order for the block kernel's queue)
//     OSA short : OSA patterns of 33-64 units, then of 1-32 units, each by text
//                 length, longest first
//     segmented : Levenshtein / Indel / LCSseq pairs whose shorter string has
//                 <= 4096 units.  The pattern is the shorter string, or the
//                 longer one when that has C-4096 units (C = chain_steps, 512
//                 by default: a 60-unit string against a 3000-unit one then
//                 runs 64 lanes x 60 steps instead of 1 lane x 3000 steps).
//                 16 ranges: pairs with texts of >= C units first (long
//                 sequential chains), then the rest;
//                 within each, lanes-per-pair classes 64, 32, ..., 1 and then
//                 one lane with a pattern of <= 32 units; within a
//                 class Levenshtein before the LCS metrics, each by text length
//                 (exact below 256 units, 32 bins per power of two above),
//                 longest first, so the lanes of a wave finish together.
//   passes   : k_bin_count   -- bin of every pair, per-workgroup histograms in
//                               LDS, each written whole (coalesced);
//              k_bin_colscan -- per bin, the running count over the workgroups
//                               (64 bins per workgroup, its 16 waves on slices
//                               of the workgroups) and the bin's total;
//              k_bin_scatter -- every workgroup scans the bin totals in LDS (the
//                               start of each bin); a pair's output position is
//                               its bin's start + the workgroups before it + its
//                               rank inside its workgroup, placed with LDS
//                               cursors; workgroup 0 writes the range bounds.
// The host never reads the plan back, so a call costs O(1) host work.
#include <algorithm>

#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

constexpr uint32_t SEG_MAX_UNITS = 4096;
constexpr uint32_t TBINS = 1024; // text-length bins
constexpr uint32_t BLOCK_BINS = 256;
constexpr uint32_t OSA_BASE = BLOCK_BINS;
constexpr uint32_t SEG_BASE = OSA_BASE + 2 * TBINS;
constexpr uint32_t NBINS = SEG_BASE + 2 * SEG_CLASSES * 2 * TBINS; // 35,072
constexpr uint32_t PLAN_THREADS = 1024;
constexpr uint32_t PLAN_MAX_GROUPS = 128;

// text length -> [0, TBINS): exact below 256, then 32 bins per power of two
__device__ __forceinline__ uint32_t text_bin(uint64_t t)
{
    const uint32_t v = (uint32_t)(t < 0xffffffffull ? t : 0xffffffffull);
    if (v < 256) return v;
    const uint32_t lg = 31 - __builtin_clz(v);
    return 256 + (lg - 8) * 32 + ((v >> (lg - 5)) & 31);
}

// cost -> [0, 256): 4 bins per power of two
__device__ __forceinline__ uint32_t cost_code(uint64_t c)
{
    if (c < 4) return (uint32_t)c;
    const uint32_t lg = 63 - __builtin_clzll(c);
    return lg * 4 + (uint32_t)((c >> (lg - 2)) & 3);
}

__device__ __forceinline__ uint32_t words_class(uint64_t m)
{
    uint32_t c = 0;
    while ((1ull << c) < (m + 63) / 64) ++c;
    return c; // log2 of lanes per pair, 0..6 for m <= 4096
}

// bin and packed item (pair | pattern-is-B << 31) of pair i
__device__ __forceinline__ uint32_t pair_bin(const uint64_t* __restrict__ aoff, const uint64_t* __restrict__ boff,
                                             const uint8_t* __restrict__ metric, uint64_t i, uint32_t chain,
                                             uint32_t* packed)
{
    const uint64_t la = aoff[i + 1] - aoff[i], lb = boff[i + 1] - boff[i];
    bool pat_b = lb < la; // the pattern is the shorter string ...
    const uint64_t s = pat_b ? lb : la, l = pat_b ? la : lb;
    const int32_t met = (int32_t)metric[i];
    uint32_t bin;
    if (met != RFA_OSA && s <= SEG_MAX_UNITS) {
        uint64_t m = s, t = l;
        if (l <= SEG_MAX_UNITS && l >= chain) {
            // ... except in the segmented kernel when the longer string has chain-4096 units
            pat_b = !pat_b;
            m = l;
            t = s;
        }
        const uint32_t cls = m <= 32 ? SEG_HALF_CLASS : 6u - words_class(m);
        const uint32_t r = (t >= chain ? 0u : SEG_CLASSES) + cls;
        bin = SEG_BASE + (2 * r + (met == RFA_LEVENSHTEIN ? 0u : 1u)) * TBINS + (TBINS - 1 - text_bin(t));
    }
    else if (met == RFA_OSA && s <= 64) {
        bin = OSA_BASE + (s > 32 ? 0u : TBINS) + (TBINS - 1 - text_bin(l));
    }
    else {
        const uint32_t c = cost_code(((s + 63) / 64) * l);
        bin = BLOCK_BINS - 1 - (c < BLOCK_BINS ? c : BLOCK_BINS - 1);
    }
    *packed = (uint32_t)i | (pat_b ? 0x80000000u : 0u);
    return bin;
}

// Workgroup g bins pairs [g * chunk, min(n, (g + 1) * chunk)); hist[g * NBINS + bin].
__global__ __launch_bounds__(PLAN_THREADS) void k_bin_count(const uint64_t* __restrict__ aoff,
                                                            const uint64_t* __restrict__ boff,
                                                            const uint8_t* __restrict__ metric, uint64_t n,
                                                            uint32_t chain, uint64_t chunk, uint32_t* __restrict__ bins,
                                                            uint32_t* __restrict__ vals, uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[NBINS];
    for (uint32_t b = threadIdx.x; b < NBINS; b += PLAN_THREADS) h[b] = 0;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = std::min<uint64_t>(n, lo + chunk);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += PLAN_THREADS) {
        uint32_t pk;
        const uint32_t b = pair_bin(aoff, boff, metric, i, chain, &pk);
        bins[i] = b;
        vals[i] = pk;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < NBINS; b += PLAN_THREADS) hist[(uint64_t)blockIdx.x * NBINS + b] = h[b];
}

// pos[g * NBINS + bin] = pairs of the bin in workgroups before g; tot[bin] = the bin's
// pairs.  A workgroup takes 64 bins (one per lane); wave w sums its slice of the
// workgroups, the waves exchange their sums in LDS, then each rewrites its slice as
// running counts.  Every global access is 64 consecutive bins of one workgroup's row.
constexpr uint32_t COLSCAN_WAVES = 16;
static_assert(NBINS % 64 == 0, "whole tiles of 64 bins");
__global__ __launch_bounds__(64 * COLSCAN_WAVES) void k_bin_colscan(const uint32_t* __restrict__ hist, uint32_t groups,
                                                                    uint32_t* __restrict__ pos,
                                                                    uint32_t* __restrict__ tot)
{
    __shared__ uint32_t part[COLSCAN_WAVES][64];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t b = (uint64_t)blockIdx.x * 64 + lane;
    const uint32_t slice = (groups + COLSCAN_WAVES - 1) / COLSCAN_WAVES;
    const uint32_t g0 = std::min(groups, w * slice), g1 = std::min(groups, g0 + slice);
    uint32_t sum = 0;
    for (uint32_t g = g0; g < g1; ++g) sum += hist[(uint64_t)g * NBINS + b];
    part[w][lane] = sum;
    __syncthreads();
    uint32_t run = 0, all = 0;
    for (uint32_t v = 0; v < COLSCAN_WAVES; ++v) {
        const uint32_t x = part[v][lane];
        run += v < w ? x : 0u;
        all += x;
    }
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t x = hist[(uint64_t)g * NBINS + b];
        pos[(uint64_t)g * NBINS + b] = run;
        run += x;
    }
    if (w == 0) tot[b] = all;
}

// The start of every bin: exclusive scan of tot, in LDS.  Thread t owns bins
// [t * BPT, (t + 1) * BPT); returns with start[bin] in s[bin].
constexpr uint32_t SCAN_BPT = (NBINS + PLAN_THREADS - 1) / PLAN_THREADS;
__device__ __forceinline__ void bin_starts(const uint32_t* __restrict__ tot, uint32_t* s, uint32_t* tmp)
{
    for (uint32_t b = threadIdx.x; b < NBINS; b += PLAN_THREADS) s[b] = tot[b];
    __syncthreads();
    const uint32_t b0 = threadIdx.x * SCAN_BPT, b1 = std::min(NBINS, b0 + SCAN_BPT);
    uint32_t v = 0;
    for (uint32_t b = b0; b < b1; ++b) v += s[b];
    uint32_t total;
    uint32_t run = block_exclusive_scan<PLAN_THREADS / 64>(v, tmp, &total);
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t x = s[b];
        s[b] = run;
        run += x;
    }
    __syncthreads();
}

// counters[0]: block queue head (0); counters[1..3]: ends of the block and OSA ranges;
// counters[4 + r]: end of segmented range r
__device__ __forceinline__ void write_bounds(const uint32_t* s, uint64_t n, uint32_t* __restrict__ counters)
{
    const uint32_t t = threadIdx.x;
    auto start_of = [&](uint32_t bin) { return bin < NBINS ? s[bin] : (uint32_t)n; };
    if (t == 0) counters[0] = 0;
    if (t == 1) counters[1] = start_of(OSA_BASE);
    if (t == 2) counters[2] = start_of(OSA_BASE + TBINS);
    if (t == 3) counters[3] = start_of(SEG_BASE);
    if (t < 2 * SEG_CLASSES) counters[4 + t] = start_of(SEG_BASE + (t + 1) * 2 * TBINS);
}

__global__ __launch_bounds__(PLAN_THREADS) void k_bin_scatter(uint64_t n, uint64_t chunk,
                                                              const uint32_t* __restrict__ bins,
                                                              const uint32_t* __restrict__ vals,
                                                              const uint32_t* __restrict__ pos,
                                                              const uint32_t* __restrict__ tot,
                                                              uint32_t* __restrict__ out,
                                                              uint32_t* __restrict__ counters)
{
    __shared__ uint32_t cur[NBINS];
    __shared__ uint32_t tmp[PLAN_THREADS / 64];
    bin_starts(tot, cur, tmp);
    if (blockIdx.x == 0) write_bounds(cur, n, counters);
    for (uint32_t b = threadIdx.x; b < NBINS; b += PLAN_THREADS) cur[b] += pos[(uint64_t)blockIdx.x * NBINS + b];
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = std::min<uint64_t>(n, lo + chunk);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += PLAN_THREADS) out[atomicAdd(&cur[bins[i]], 1u)] = vals[i];
}

static uint32_t plan_groups(uint64_t n)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(PLAN_MAX_GROUPS, (n + 8191) / 8192));
}

// temp: hist and pos (NBINS per workgroup each) and the bin totals
size_t plan_temp_bytes(uint64_t n)
{
    const size_t cells = (size_t)NBINS * plan_groups(n);
    return (2 * cells + NBINS) * sizeof(uint32_t) + 256;
}

hipError_t launch_plan(const uint64_t* aoff, const uint64_t* boff, const uint8_t* metric, uint64_t n,
                       uint32_t chain_steps, uint32_t* bins, uint32_t* vals_in, uint32_t* vals_out, void* temp,
                       size_t temp_bytes, uint32_t* counters, hipStream_t s)
{
    const uint32_t groups = plan_groups(n);
    const uint64_t chunk = (n + groups - 1) / groups;
    const size_t cells = (size_t)NBINS * groups;
    if (temp_bytes < plan_temp_bytes(n)) return hipErrorInvalidValue;
    uint32_t* hist = (uint32_t*)temp;
    uint32_t* pos = hist + cells;
    uint32_t* tot = pos + cells;
    hipLaunchKernelGGL(k_bin_count, dim3(groups), dim3(PLAN_THREADS), 0, s, aoff, boff, metric, n, chain_steps, chunk,
                       bins, vals_in, hist);
    hipLaunchKernelGGL(k_bin_colscan, dim3(NBINS / 64), dim3(64 * COLSCAN_WAVES), 0, s, hist, groups, pos, tot);
    hipLaunchKernelGGL(k_bin_scatter, dim3(groups), dim3(PLAN_THREADS), 0, s, n, chunk, bins, vals_in, pos, tot,
                       vals_out, counters);
    return hipGetLastError();
}

} // namespace rfa
