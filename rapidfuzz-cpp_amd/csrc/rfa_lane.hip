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
    for (uint64_t item = g; item < a.nitems; item += nlanes) {
        const uint32_t ia = a.item_a ? a.item_a[item] : (uint32_t)item;
        const uint32_t ib = a.item_b ? a.item_b[item] : (uint32_t)item;
        const bool swap = a.item_swap ? a.item_swap[item] != 0 : false;
        const uint64_t oa = a.aoff[ia], ob = a.boff[ib];
        const uint32_t la = (uint32_t)(a.aoff[ia + 1] - oa), lb = (uint32_t)(a.boff[ib + 1] - ob);
        const CharT* __restrict__ p = swap ? B + ob : A + oa;
        const CharT* __restrict__ t = swap ? A + oa : B + ob;
        const uint32_t m = swap ? lb : la, n = swap ? la : lb;
        ScoreSpec spec = a.spec;
        if (a.item_metric) {
            spec.metric = a.item_metric[item];
        }
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
        finalize(spec, la, lb, res, a.out, a.out_index ? a.out_index[item] : item);
    }
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
