// rfa_lane.hip -- one pair per lane.
//
// Each lane owns an independent (pattern, text) pair and walks the text one
// code unit at a time, updating the pattern's words one after another with
// the carries passed in registers (the reference's sequential block
// recurrences: Levenshtein_impl.hpp:689-743, LCSseq_impl.hpp:219-252 and
// 286-314, OSA_impl.hpp:199-243).  The pattern-match vectors are kept the
// way the reference keeps them for wide code units: one 128-slot
// open-addressing map per 64-unit block (PatternMatchVector.hpp:17-60,
// 136-221), here in a per-lane arena in global memory.  Slots carry a
// generation tag, so building the map for the next pair needs no clearing.
//
// MAXW = 1 is the short-pattern bucket of the bucketed pairs schedule (the
// words live in registers); MAXW = 64 is the plain one-pair-per-lane kernel
// used as the baseline schedule (its words live in private memory).
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"

namespace rfa {

constexpr uint32_t MAP_SLOTS = 128;

struct MapEntry {
    uint64_t tag; // (generation << 32) | symbol
    uint64_t mask;
};

__device__ __forceinline__ uint64_t map_get(const MapEntry* __restrict__ map, uint32_t c, uint32_t gen)
{
    const uint64_t want = ((uint64_t)gen << 32) | c;
    uint32_t h = (c * 0x9E3779B1u) >> 25; // 7 bits
    for (uint32_t probe = 0; probe < MAP_SLOTS; ++probe) {
        const uint64_t t = map[h].tag;
        if (t == want) return map[h].mask;
        if ((uint32_t)(t >> 32) != gen) return 0;
        h = (h + 1) & (MAP_SLOTS - 1);
    }
    return 0;
}

__device__ __forceinline__ void map_or(MapEntry* __restrict__ map, uint32_t c, uint32_t gen, uint64_t bit)
{
    const uint64_t want = ((uint64_t)gen << 32) | c;
    uint32_t h = (c * 0x9E3779B1u) >> 25;
    for (;;) {
        const uint64_t t = map[h].tag;
        if (t == want) {
            map[h].mask |= bit;
            return;
        }
        if ((uint32_t)(t >> 32) != gen) {
            map[h].tag = want;
            map[h].mask = bit;
            return;
        }
        h = (h + 1) & (MAP_SLOTS - 1);
    }
}

template <int ALGO, typename CharT, int MAXW>
__global__ __launch_bounds__(256) void k_lane(LaneArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nlanes = (uint64_t)gridDim.x * blockDim.x;
    MapEntry* __restrict__ maps = (MapEntry*)(a.arena + g * a.arena_words);
    uint32_t gen = a.gen_base;
    // per-lane word state (private memory for MAXW > 1)
    uint64_t VP[MAXW], VN[MAXW], D0s[MAXW], PMo[MAXW];
    const uint64_t begin = a.begin_dev ? *a.begin_dev : 0;
    for (uint64_t item = begin + g; item < a.nitems; item += nlanes) {
        uint32_t ia, ib;
        bool swap;
        uint64_t oidx;
        if (a.item_packed) {
            const uint32_t pk = a.item_packed[item];
            ia = ib = pk & 0x7fffffffu;
            swap = (pk >> 31) != 0;
            oidx = ia;
        }
        else {
            ia = a.item_a ? a.item_a[item] : (uint32_t)item;
            ib = a.item_b ? a.item_b[item] : (uint32_t)item;
            swap = a.item_swap ? a.item_swap[item] != 0 : false;
            oidx = a.out_index ? a.out_index[item] : item;
        }
        const uint64_t oa = a.aoff[ia], ob = a.boff[ib];
        const uint32_t la = (uint32_t)(a.aoff[ia + 1] - oa), lb = (uint32_t)(a.boff[ib + 1] - ob);
        const CharT* __restrict__ p = swap ? B + ob : A + oa;
        const CharT* __restrict__ t = swap ? A + oa : B + ob;
        const uint32_t m = swap ? lb : la, n = swap ? la : lb;
        ScoreSpec spec = a.spec;
        if (a.pair_metric)
            spec.metric = a.pair_metric[ia];
        else if (a.item_metric)
            spec.metric = a.item_metric[item];
        const int algo = (spec.metric == RFA_LEVENSHTEIN) ? ALGO_LEV : (spec.metric == RFA_OSA ? ALGO_OSA : ALGO_LCS);
        RawResult res{0, 0, 0};
        if (m == 0) {
            res.v = (algo == ALGO_LCS) ? 0 : n;
        }
        else if ((m + 63) / 64 <= (uint32_t)MAXW) { // the host routes longer patterns elsewhere
            const uint32_t W = (m + 63) / 64;
            // build: one map per 64-unit block, fresh generation per pair
            ++gen;
            for (uint32_t i = 0; i < m; ++i) map_or(maps + (i >> 6) * MAP_SLOTS, (uint32_t)p[i], gen, 1ull << (i & 63));
            for (uint32_t w = 0; w < W && w < (uint32_t)MAXW; ++w) {
                VP[w] = ~0ull;
                VN[w] = 0;
                D0s[w] = 0;
                PMo[w] = 0;
            }
            const uint64_t lastmask = (m & 63) ? ((1ull << (m & 63)) - 1) : ~0ull;
            for (uint32_t j = 0; j < n; ++j) {
                const uint32_t ch = (uint32_t)t[j];
                if (algo == ALGO_LCS) {
                    uint64_t carry = 0;
                    for (uint32_t w = 0; w < W; ++w) {
                        const uint64_t X = map_get(maps + w * MAP_SLOTS, ch, gen);
                        const uint64_t S = VP[w];
                        const uint64_t u = S & X;
                        const uint64_t x = S + u;
                        const uint64_t c1 = x < S;
                        const uint64_t x2 = x + carry;
                        const uint64_t c2 = x2 < x;
                        carry = c1 | c2;
                        VP[w] = x2 | (S - u);
                    }
                }
                else {
                    uint64_t hp_c = 1, hn_c = 0, d0_last = 0, pm_last = 0;
                    for (uint32_t w = 0; w < W; ++w) {
                        uint64_t X = map_get(maps + w * MAP_SLOTS, ch, gen);
                        const uint64_t vp = VP[w], vn = VN[w];
                        uint64_t TR = 0;
                        if (algo == ALGO_OSA) {
                            TR = (((~D0s[w]) & X) << 1 | (((~d0_last) & pm_last) >> 63)) & PMo[w];
                            d0_last = D0s[w];
                            pm_last = X;
                        }
                        const uint64_t Xc = X | hn_c;
                        const uint64_t D0 = (((Xc & vp) + vp) ^ vp) | Xc | vn | TR;
                        uint64_t HP = vn | ~(D0 | vp);
                        uint64_t HN = D0 & vp;
                        const uint64_t hp_o = HP >> 63, hn_o = HN >> 63;
                        HP = (HP << 1) | hp_c;
                        HN = (HN << 1) | hn_c;
                        VP[w] = HN | ~(D0 | HP);
                        VN[w] = HP & D0;
                        if (algo == ALGO_OSA) {
                            D0s[w] = D0;
                            PMo[w] = X;
                        }
                        hp_c = hp_o;
                        hn_c = hn_o;
                    }
                }
            }
            uint64_t acc = 0;
            if (algo == ALGO_LCS) {
                for (uint32_t w = 0; w < W; ++w) acc += __popcll(~VP[w] & (w + 1 == W ? lastmask : ~0ull));
                res.v = acc;
            }
            else {
                int64_t d = n;
                for (uint32_t w = 0; w < W; ++w) {
                    const uint64_t M = (w + 1 == W) ? lastmask : ~0ull;
                    d += (int64_t)__popcll(VP[w] & M) - (int64_t)__popcll(VN[w] & M);
                }
                res.v = (uint64_t)d;
            }
        }
        finalize(spec, la, lb, res, a.out, oidx);
    }
}

// ---- short patterns (<= 64 units): per-lane symbol tables in LDS ------------
//
// The bucketed schedule sends pairs whose shorter string has at most 64 units
// here.  Each lane keeps a 64-slot open-addressing table (symbol -> match
// mask) for its pattern in LDS, slot s of lane l at [s * 64 + l]; a slot is
// free when its mask is zero, and after a pair only the occupied slots
// (tracked in a 64-bit register) are cleared.  W is 32 bits for patterns of at
// most 32 units and 64 bits otherwise; the planner keeps the two groups (and
// the metrics) in separate ranges sorted by text length, so the lanes of a
// wave run the same recurrence for about the same number of steps.
constexpr uint32_t SHORT_SLOTS = 64;

template <typename CharT, typename W>
__global__ __launch_bounds__(64) void k_lane_short(LaneArgs a)
{
    __shared__ uint32_t keys[SHORT_SLOTS * 64];
    __shared__ W masks[SHORT_SLOTS * 64];
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x;
    for (uint32_t s = 0; s < SHORT_SLOTS; ++s) masks[s * 64 + lane] = 0;
    const uint64_t begin = a.begin_dev ? *a.begin_dev : 0;
    const uint64_t end = a.end_dev ? *a.end_dev : a.nitems;
    const uint64_t nl = (uint64_t)gridDim.x * 64;
    for (uint64_t item = begin + (uint64_t)blockIdx.x * 64 + lane; item < end; item += nl) {
        const uint32_t pk = a.item_packed[item];
        const uint32_t pair = pk & 0x7fffffffu;
        const bool swap = (pk >> 31) != 0;
        const uint64_t oa = a.aoff[pair], ob = a.boff[pair];
        const uint32_t la = (uint32_t)(a.aoff[pair + 1] - oa), lb = (uint32_t)(a.boff[pair + 1] - ob);
        const CharT* __restrict__ p = swap ? B + ob : A + oa;
        const CharT* __restrict__ t = swap ? A + oa : B + ob;
        const uint32_t m = swap ? lb : la, n = swap ? la : lb;
        ScoreSpec spec = a.spec;
        spec.metric = a.pair_metric[pair];
        const int algo = (spec.metric == RFA_LEVENSHTEIN) ? ALGO_LEV : (spec.metric == RFA_OSA ? ALGO_OSA : ALGO_LCS);
        RawResult res{0, 0, 0};
        if (m == 0) {
            res.v = (algo == ALGO_LCS) ? 0 : n;
        }
        else {
            uint64_t occ = 0;
            uint32_t pat[64];
#pragma unroll
            for (uint32_t i = 0; i < 64; ++i) pat[i] = (uint32_t)p[min(i, m - 1)]; // independent loads
#pragma unroll
            for (uint32_t i = 0; i < 64; ++i) {
                if (i < m) {
                    const uint32_t c = pat[i];
                    uint32_t h = (c * 0x9E3779B1u) >> 26;
                    while (masks[h * 64 + lane] != 0 && keys[h * 64 + lane] != c) h = (h + 1) & (SHORT_SLOTS - 1);
                    keys[h * 64 + lane] = c;
                    masks[h * 64 + lane] |= (W)1 << i;
                    occ |= 1ull << h;
                }
            }
            uint32_t VPl = ~0u, VPh = ~0u, VNl = 0, VNh = 0, D0l = 0, D0h = 0, PMl = 0, PMh = 0;
            // the text is read CHUNK units ahead: one batch of independent loads is in
            // flight while the previous batch is consumed
            constexpr uint32_t CHUNK = 8;
            uint32_t nxt[CHUNK];
            // loads are unconditional (index clamped to the last unit) so the wait
            // counter can retire them in order instead of draining all of them
            const uint32_t last = n - 1;
#pragma unroll
            for (uint32_t k = 0; k < CHUNK; ++k) nxt[k] = (uint32_t)t[min(k, last)];
            for (uint32_t j0 = 0; j0 < n; j0 += CHUNK) {
                uint32_t cur[CHUNK];
#pragma unroll
                for (uint32_t k = 0; k < CHUNK; ++k) {
                    cur[k] = nxt[k];
                    nxt[k] = (uint32_t)t[min(j0 + CHUNK + k, last)];
                }
                const uint32_t cnt = min(CHUNK, n - j0);
#pragma unroll
                for (uint32_t k = 0; k < CHUNK; ++k) {
                if (k < cnt) {
                const uint32_t c = cur[k];
                uint32_t h = (c * 0x9E3779B1u) >> 26;
                W X = masks[h * 64 + lane];
                uint32_t key = keys[h * 64 + lane];
                while (X != 0 && key != c) {
                    h = (h + 1) & (SHORT_SLOTS - 1);
                    X = masks[h * 64 + lane];
                    key = keys[h * 64 + lane];
                }
                if constexpr (sizeof(W) == 4) {
                    if (algo == ALGO_LEV)
                        lev_step((uint32_t)X, VPl, VNl);
                    else if (algo == ALGO_OSA)
                        osa_step((uint32_t)X, VPl, VNl, D0l, PMl);
                    else
                        lcs_step((uint32_t)X, VPl);
                }
                else {
                    if (algo == ALGO_LEV)
                        lev_step((uint64_t)X, VPl, VPh, VNl, VNh);
                    else if (algo == ALGO_OSA)
                        osa_step((uint64_t)X, VPl, VPh, VNl, VNh, D0l, D0h, PMl, PMh);
                    else
                        lcs_step((uint64_t)X, VPl, VPh);
                }
                }
                }
            }
            while (occ) {
                const uint32_t h = __builtin_ctzll(occ);
                occ &= occ - 1;
                masks[h * 64 + lane] = 0;
            }
            const W VP = (sizeof(W) == 8) ? (W)(((uint64_t)VPh << 32) | VPl) : (W)VPl;
            const W VN = (sizeof(W) == 8) ? (W)(((uint64_t)VNh << 32) | VNl) : (W)VNl;
            const W M = (m >= sizeof(W) * 8) ? ~(W)0 : (((W)1 << m) - 1);
            if (algo == ALGO_LCS)
                res.v = (uint64_t)popc<W>(~VP & M);
            else
                res.v = (uint64_t)n + popc<W>(VP & M) - popc<W>(VN & M);
        }
        finalize(spec, la, lb, res, a.out, pair);
    }
}

hipError_t launch_lane_short(int char_width, int word_bits, const LaneArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (word_bits == 32) {
        if (char_width == 1)
            hipLaunchKernelGGL((k_lane_short<uint8_t, uint32_t>), dim3(nblocks), dim3(64), 0, s, a);
        else if (char_width == 2)
            hipLaunchKernelGGL((k_lane_short<uint16_t, uint32_t>), dim3(nblocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_lane_short<uint32_t, uint32_t>), dim3(nblocks), dim3(64), 0, s, a);
    }
    else {
        if (char_width == 1)
            hipLaunchKernelGGL((k_lane_short<uint8_t, uint64_t>), dim3(nblocks), dim3(64), 0, s, a);
        else if (char_width == 2)
            hipLaunchKernelGGL((k_lane_short<uint16_t, uint64_t>), dim3(nblocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_lane_short<uint32_t, uint64_t>), dim3(nblocks), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

size_t lane_arena_words(uint32_t maxw) { return (size_t)maxw * MAP_SLOTS * 2; }

template <typename CharT, int MAXW>
static void launch_lane_w(const LaneArgs& a, uint32_t nblocks, hipStream_t s)
{
    // the recurrence is chosen per item from the metric, so one instantiation serves all
    hipLaunchKernelGGL((k_lane<0, CharT, MAXW>), dim3(nblocks), dim3(256), 0, s, a);
}

hipError_t launch_lane(int char_width, int maxw, const LaneArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (maxw <= 1) {
        if (char_width == 1)
            launch_lane_w<uint8_t, 1>(a, nblocks, s);
        else if (char_width == 2)
            launch_lane_w<uint16_t, 1>(a, nblocks, s);
        else
            launch_lane_w<uint32_t, 1>(a, nblocks, s);
    }
    else {
        if (char_width == 1)
            launch_lane_w<uint8_t, 64>(a, nblocks, s);
        else if (char_width == 2)
            launch_lane_w<uint16_t, 64>(a, nblocks, s);
        else
            launch_lane_w<uint32_t, 64>(a, nblocks, s);
    }
    return hipGetLastError();
}

} // namespace rfa
