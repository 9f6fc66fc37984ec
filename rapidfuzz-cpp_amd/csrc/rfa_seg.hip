// rfa_seg.hip -- segmented pairs kernel: independent pairs whose pattern has at
// most 4096 units and whose metric runs the Levenshtein or the LCS recurrence
// (Levenshtein, Indel, LCSseq).  The planner picks the pattern: the shorter
// string, or the longer one when that has 512-4096 units, which turns one long
// sequential chain (1 lane x 3000 text steps) into a wide short one (47 lanes x
// 60 steps) at no more lane-steps.
//
// Layout: a pair takes a segment of SEG lanes (SEG = the power of two >= its
// word count, 1..64); lane w of the segment holds pattern word w, so a wave
// runs 64 / SEG pairs at once.  The planner sorts pairs by (SEG class, metric,
// text length), so a wave's pairs share their recurrence and trip count.
//
// Match masks without a symbol map: each lane keeps the bit-planes of its 64
// pattern units, plane_b bit i = bit b of unit i, and for text symbol c
//   PM(c) = AND_b (plane_b XNOR bit_b(c))   over the NB low symbol bits
// (NB >= bit length of the largest symbol of the launch; a text symbol with a
// bit set at or above NB matches nothing).  That is NB three-input bit ops per
// 32-bit half word and no LDS, so occupancy is set by registers alone.  The
// planes of 32 units are one transpose of the 32 x NB bit matrix (units as
// rows): a byte-permute stage, then masked swaps of 4-, 2- and 1-bit blocks,
// about 4 VALU per unit.
//
// Carries between the words of a pair use the carry-lookahead of the block
// kernel (generate / propagate ballots resolved by one 64-bit add), with the
// last lane of every segment masked out of the add so no carry crosses into
// the next pair, and the HP carry-in forced to 1 at segment starts
// (Levenshtein_impl.hpp:700-736, LCSseq_impl.hpp:286-314).
//
// Work: group g of range r holds items [begin_r + g * PPW, +PPW) of the sorted
// list: long-text pairs first, then widest class first and longest texts first
// within a class.  The grid
// is exactly the resident waves, and wave w takes groups w, w + W, w + 2W, ...
// (round-robin over a longest-first list).  No shared queue: a device-wide
// atomic per group costs more than a short group's whole scan.
#include <algorithm>
#include <type_traits>

#include "rfa_block_impl.hpp"
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"

namespace rfa {

// bit i set for every lane i that starts a segment of 2^lseg lanes (wave-uniform)
__device__ __forceinline__ uint64_t seg_first_mask(uint32_t lseg)
{
    uint64_t m = 1;
    for (uint32_t s = 1u << lseg; s < 64; s <<= 1) m |= m << s;
    return m;
}

constexpr uint32_t T_AND_XNOR = (TA & ~(TB ^ TC)) & 0xFF; // a & ~(b ^ c)

// Bit-planes of 32 pattern units: pl[b] bit i = bit b of unit i0 + i, from 4 loads of
// 8 units issued together (a chunk at or past cnt re-reads the pattern's start, so every
// load stays inside the set).  Units past cnt are whatever the loads return; their bits
// reach only plane positions >= cnt - i0, which the caller masks off.  The transpose runs
// in place over r[] (rows = units, columns = symbol bits): blocks of 16 and 8 bits trade
// places by byte permutes (or a shifted OR where the partner block is known zero: symbols
// below 2^NB), then 4-, 2- and 1-bit blocks by shift + bit-field insert.
template <typename CharT, uint32_t NB>
__device__ __forceinline__ void seg_planes(const CharT* __restrict__ p, uint32_t cnt, uint32_t i0, uint32_t (&pl)[NB])
{
    uint32_t r[32];
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) load_units<CharT, 8>(i0 + 8 * c < cnt ? p + i0 + 8 * c : p, r + 8 * c);
    if constexpr (NB <= 16) {
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) r[i] |= r[i + 16] << 16;
    }
    else {
#pragma unroll
        for (uint32_t i = 0; i < 16; ++i) {
            const uint32_t a = r[i], b = r[i + 16];
            r[i + 16] = __builtin_amdgcn_perm(a, b, 0x03020706u);
            r[i] = __builtin_amdgcn_perm(a, b, 0x01000504u);
        }
    }
    constexpr uint32_t R16 = NB <= 16 ? 16 : 32;
    if constexpr (NB <= 8) {
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i) r[i] |= r[i + 8] << 8;
    }
    else {
#pragma unroll
        for (uint32_t base = 0; base < R16; base += 16)
#pragma unroll
            for (uint32_t i = base; i < base + 8; ++i) {
                const uint32_t a = r[i], b = r[i + 8];
                r[i + 8] = __builtin_amdgcn_perm(a, b, 0x03070105u);
                r[i] = __builtin_amdgcn_perm(a, b, 0x02060004u);
            }
    }
    constexpr uint32_t R = NB <= 8 ? 8 : R16;
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
        const uint32_t sh = 4u >> k;
        const uint32_t m = k == 0 ? 0x0F0F0F0Fu : (k == 1 ? 0x33333333u : 0x55555555u);
#pragma unroll
        for (uint32_t i = 0; i < R; ++i)
            if ((i & sh) == 0) {
                const uint32_t a = r[i], b = r[i + sh];
                r[i + sh] = (m & (a >> sh)) | (~m & b);
                r[i] = (~m & (b << sh)) | (m & a);
            }
    }
#pragma unroll
    for (uint32_t b = 0; b < NB; ++b) pl[b] = r[b];
}

// Recurrence over the text for one word per lane; returns the lane's share of
// the pair's raw value (Levenshtein: sum of vertical deltas of the last
// column; LCS: matched rows).
// MULTI: segments of 2^lseg >= 2 lanes (the carries cross lanes); otherwise one lane
// per pair.  All multi-lane classes share one body with their masks as wave-uniform
// values, which keeps the kernel's code within the instruction cache.  HALF (one lane,
// patterns of at most 32 units): the word, the match mask and the recurrence are 32-bit.
template <bool MULTI, bool HALF, typename CharT, uint32_t NB>
__device__ __forceinline__ int32_t seg_run(const CharT* __restrict__ p, const CharT* __restrict__ t, uint32_t cnt,
                                           uint32_t n, bool wact, bool lev, uint32_t nmax, uint32_t lane,
                                           uint32_t lseg)
{
    const uint64_t FIRSTM = MULTI ? seg_first_mask(lseg) : ~0ull;
    const uint64_t LASTM = MULTI ? FIRSTM << ((1u << lseg) - 1) : ~0ull;
    // bit-planes of this word's units (lo: units 0-31, hi: units 32-63); units past
    // cnt only set bits that vml / vmh mask off
    static_assert(!(MULTI && HALF), "half words run one lane per pair");
    using Wd = std::conditional_t<HALF, uint32_t, uint64_t>;
    constexpr uint32_t CH = 8;
    uint32_t pl[NB], ph[NB];
    seg_planes<CharT, NB>(p, cnt, 0u, pl);
    if (!HALF && ballot(cnt > 32) != 0) {
        seg_planes<CharT, NB>(p, cnt, 32u, ph);
    }
    else {
#pragma unroll
        for (uint32_t b = 0; b < NB; ++b) ph[b] = 0;
    }
    const uint32_t vml = cnt >= 32 ? ~0u : ((1u << cnt) - 1u);
    const uint32_t vmh = cnt >= 64 ? ~0u : (cnt > 32 ? ((1u << (cnt - 32)) - 1u) : 0u);

    const bool any_lev = ballot(wact && lev) != 0, any_lcs = ballot(wact && !lev) != 0;
    Wd VP = ~(Wd)0, VN = 0; // LCS keeps S in VP

    // one text position: column j, text unit c
    auto step = [&](uint32_t j, uint32_t c) {
        const bool act = wact && j < n;
        uint32_t xl = vml, xh = vmh;
        if constexpr (NB < 32) {
            if ((c >> NB) != 0) xl = xh = 0;
        }
#pragma unroll
        for (uint32_t b = 0; b < NB; ++b) {
            const uint32_t sb = (uint32_t)(-(int32_t)((c >> b) & 1u)); // all ones iff bit b of c
            xl = bop3<T_AND_XNOR>(xl, pl[b], sb);
            if constexpr (!HALF) xh = bop3<T_AND_XNOR>(xh, ph[b], sb);
        }
        Wd X;
        if constexpr (HALF)
            X = xl;
        else
            X = ((uint64_t)xh << 32) | xl;
        if (any_lev) {
            const bool al = act && lev;
            const Wd vp = VP, vn = VN;
            Wd D0, HP, HN;
            if constexpr (!MULTI) {
                D0 = (((X & vp) + vp) ^ vp) | X | vn;
                HP = vn | ~(D0 | vp);
                HN = D0 & vp;
                HP = (HP << 1) | 1u;
                HN = HN << 1;
            }
            else {
                const uint64_t X1 = X | 1;
                const uint64_t d0a = (((X & vp) + vp) ^ vp) | X | vn;
                const uint64_t d0b = (((X1 & vp) + vp) ^ vp) | X1 | vn;
                // outgoing HN carry for incoming carry 0 / 1 (monotone in it)
                const bool ga = al && ((d0a & vp) >> 63);
                const bool gb = al && ((d0b & vp) >> 63);
                const uint64_t G = ballot(ga) & ~LASTM, Pp = ballot(gb && !ga) & ~LASTM;
                const uint32_t hin = lane_bit(((G | Pp) + G) ^ Pp, lane);
                D0 = hin ? d0b : d0a;
                HP = vn | ~(D0 | vp);
                HN = D0 & vp;
                const uint64_t HPm = ballot(al && (HP >> 63));
                const uint32_t hpin = lane_bit((HPm << 1) | FIRSTM, lane);
                HP = (HP << 1) | hpin;
                HN = (HN << 1) | hin;
            }
            if (al) {
                VP = HN | ~(D0 | HP);
                VN = HP & D0;
            }
        }
        if (any_lcs) {
            const bool ac = act && !lev;
            const Wd Sv = VP;
            const Wd u = Sv & X;
            Wd x = Sv + u;
            if constexpr (MULTI) {
                const bool gg = ac && (x < Sv);
                const bool pp = ac && !gg && (x == ~0ull);
                const uint64_t G = ballot(gg) & ~LASTM, Pp = ballot(pp) & ~LASTM;
                x += lane_bit(((G | Pp) + G) ^ Pp, lane);
            }
            if (ac) VP = x | (Sv - u);
        }
    };
    auto chunk = [&](const uint32_t* u, uint32_t j0) {
#pragma unroll
        for (uint32_t k = 0; k < CH; ++k) {
            if (j0 + k >= nmax) break;
            step(j0 + k, u[k]);
        }
    };
    // text units CH at a time in two register buffers: one chunk is loaded while
    // the other runs (fixed roles, so no register copies between chunks).  Every
    // lane issues every load -- lanes with nothing left re-read the start of their
    // text -- so the loads stay unconditional and their waits exact.
    const bool tl = wact && n > 0;
    uint32_t ua[CH], ub[CH];
    load_units<CharT, CH>(t, ua);
    for (uint32_t j0 = 0; j0 < nmax; j0 += 2 * CH) {
        load_units<CharT, CH>(tl && j0 + CH < n ? t + j0 + CH : t, ub);
        chunk(ua, j0);
        if (j0 + CH >= nmax) break;
        load_units<CharT, CH>(tl && j0 + 2 * CH < n ? t + j0 + 2 * CH : t, ua);
        chunk(ub, j0 + CH);
    }
    if (!wact) return 0;
    if constexpr (HALF) {
        if (lev) return __popc(VP & vml) - __popc(VN & vml);
        return __popc(~VP & vml);
    }
    else {
        const uint64_t M = ((uint64_t)vmh << 32) | vml;
        if (lev) return __popcll(VP & M) - __popcll(VN & M);
        return __popcll(~VP & M);
    }
}

// One group of 64 >> lseg pairs (segments of 2^lseg lanes), items [first, min(first + pairs, end)).
// INTK: integer result kinds (distance, similarity) with the integer epilogue.
template <bool MULTI, bool HALF, typename CharT, uint32_t NB, bool INTK>
__device__ __forceinline__ void seg_group(const SegArgs& a, uint32_t first, uint32_t end, uint32_t lane, uint32_t lseg)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t slot = lane >> lseg, w = lane & ((1u << lseg) - 1);
    const uint32_t item = first + slot;
    const bool valid = item < end;
    uint32_t pair = 0, la = 0, lb = 0, m = 0, n = 0;
    const CharT* p = A;
    const CharT* t = B;
    ScoreSpec spec = a.spec;
    if (valid) {
        const uint32_t pk = a.item_packed[item];
        pair = pk & 0x7fffffffu;
        const bool swap = (pk >> 31) != 0;
        const uint64_t oa = a.aoff[pair], ob = a.boff[pair];
        la = (uint32_t)(a.aoff[pair + 1] - oa);
        lb = (uint32_t)(a.boff[pair + 1] - ob);
        p = swap ? B + ob : A + oa;
        t = swap ? A + oa : B + ob;
        m = swap ? lb : la;
        n = swap ? la : lb;
        spec.metric = a.pair_metric[pair];
        if (spec.kind == RFA_DISTANCE && spec.metric != RFA_LCSSEQ) spec.band_k = spec.cutoff_u;
    }
    const bool lev = spec.metric == RFA_LEVENSHTEIN;
    RawResult res{0, 0, 0};
    bool done = !valid;
    if (valid && length_filter(spec, lev ? ALGO_LEV : ALGO_LCS, la, lb, &res)) done = true;
    const uint32_t W = (m + 63) / 64;
    const bool wact = !done && w < W;
    const uint32_t cnt = wact ? min(64u, m - w * 64) : 0u; // pattern units of this word
    uint32_t nmax = wact ? n : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nmax = max(nmax, (uint32_t)__shfl_xor(nmax, off, 64));
    nmax = uniform(nmax);
    int32_t v = seg_run<MULTI, HALF, CharT, NB>(p + (wact ? w * 64 : 0u), t, cnt, n, wact, lev, nmax, lane, lseg);
    if constexpr (MULTI)
        for (uint32_t off = 1u << (lseg - 1); off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (valid && w == 0) {
        if (!done) res.v = lev ? (uint64_t)((int64_t)n + v) : (uint64_t)v;
        if constexpr (INTK)
            ((uint64_t*)a.out)[pair] = finalize_int(spec, la, lb, res.v);
        else
            finalize(spec, la, lb, res, a.out, pair);
    }
}

// Ranges r = 0..15: counters[3 + r] .. counters[4 + r], class r % 8; ranges
// r < 8 hold the long-text pairs, whose waves run at raised issue priority so the
// longest sequential chains are not slowed by the many short groups around them.
template <typename CharT, uint32_t NB, bool INTK>
__global__ __launch_bounds__(256) void k_seg(SegArgs a)
{
    constexpr uint32_t NR = 2 * SEG_CLASSES;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t gend[NR], ibeg[NR];
    uint32_t acc = 0;
#pragma unroll
    for (uint32_t r = 0; r < NR; ++r) {
        const uint32_t b = uniform(a.counters[3 + r]), e = uniform(a.counters[4 + r]);
        const uint32_t k = r % SEG_CLASSES;
        const uint32_t ppw = k == SEG_HALF_CLASS ? 64u : 1u << k; // pairs per wave: 64 / SEG
        ibeg[r] = b;
        acc += e > b ? (e - b + ppw - 1) / ppw : 0u;
        gend[r] = acc;
    }
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64);
    for (uint32_t g = blockIdx.x * (blockDim.x / 64) + uniform(threadIdx.x / 64); g < acc; g += nwaves) {
        uint32_t r = 0;
        while (r < NR - 1 && g >= gend[r]) ++r;
        const uint32_t g0 = r ? gend[r - 1] : 0u;
        const uint32_t k = r % SEG_CLASSES;
        const uint32_t first = ibeg[r] + (g - g0) * (k == SEG_HALF_CLASS ? 64u : 1u << k);
        const uint32_t end = uniform(a.counters[4 + r]);
        if (r < SEG_CLASSES)
            __builtin_amdgcn_s_setprio(2);
        else
            __builtin_amdgcn_s_setprio(0);
        // class k < 7: segments of 64 >> k lanes; class 7: one lane, 32-bit words
        if (k == SEG_HALF_CLASS)
            seg_group<false, true, CharT, NB, INTK>(a, first, end, lane, 0u);
        else if (k == SEG_HALF_CLASS - 1)
            seg_group<false, false, CharT, NB, INTK>(a, first, end, lane, 0u);
        else
            seg_group<true, false, CharT, NB, INTK>(a, first, end, lane, 6u - k);
    }
}

uint32_t seg_symbol_bits(int char_width, uint64_t max_symbol)
{
    uint32_t bits = 1;
    while (bits < 32 && (max_symbol >> bits) != 0) ++bits;
    if (char_width == 1 || bits <= 8) return 8;
    if (char_width == 2 || bits <= 16) return 16;
    return bits <= 24 ? 24 : 32;
}

// resident workgroups per CU of one instantiation (queried once)
template <typename CharT, uint32_t NB, bool INTK>
static uint32_t seg_blocks_per_cu()
{
    static const uint32_t v = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_seg<CharT, NB, INTK>, 256, 0) != hipSuccess || b < 1) {
            (void)hipGetLastError();
            b = 1;
        }
        return (uint32_t)b;
    }();
    return v;
}

template <typename CharT, uint32_t NB, bool INTK>
static void launch_one_seg(const SegArgs& a, uint32_t num_cus, uint32_t max_blocks, hipStream_t s)
{
    uint32_t blocks = num_cus * seg_blocks_per_cu<CharT, NB, INTK>();
    blocks = std::max(1u, std::min(blocks, max_blocks));
    hipLaunchKernelGGL((k_seg<CharT, NB, INTK>), dim3(blocks), dim3(256), 0, s, a);
}

template <bool INTK>
static void launch_seg_k(int char_width, uint32_t nb, const SegArgs& a, uint32_t num_cus, uint32_t max_blocks,
                         hipStream_t s)
{
    if (char_width == 1)
        launch_one_seg<uint8_t, 8, INTK>(a, num_cus, max_blocks, s);
    else if (char_width == 2) {
        if (nb <= 8)
            launch_one_seg<uint16_t, 8, INTK>(a, num_cus, max_blocks, s);
        else
            launch_one_seg<uint16_t, 16, INTK>(a, num_cus, max_blocks, s);
    }
    else {
        if (nb <= 8)
            launch_one_seg<uint32_t, 8, INTK>(a, num_cus, max_blocks, s);
        else if (nb <= 16)
            launch_one_seg<uint32_t, 16, INTK>(a, num_cus, max_blocks, s);
        else if (nb <= 24)
            launch_one_seg<uint32_t, 24, INTK>(a, num_cus, max_blocks, s);
        else
            launch_one_seg<uint32_t, 32, INTK>(a, num_cus, max_blocks, s);
    }
}

hipError_t launch_seg(int char_width, uint32_t nb, const SegArgs& a, uint32_t num_cus, uint32_t max_blocks,
                      hipStream_t s)
{
    if (a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY)
        launch_seg_k<true>(char_width, nb, a, num_cus, max_blocks, s);
    else
        launch_seg_k<false>(char_width, nb, a, num_cus, max_blocks, s);
    return hipGetLastError();
}

} // namespace rfa
