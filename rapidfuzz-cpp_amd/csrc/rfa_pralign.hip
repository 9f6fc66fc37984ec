// rfa_pralign.hip -- fuzz::partial_ratio_alignment of pairs (fuzz.hpp:129-135,
// fuzz_impl.hpp:68-230): the best partial ratio and the windows it was found at.
//
// One wave per pair.  The needle N (the shorter string, l1 units; the first string
// when the lengths are equal) is searched in the haystack H (l2 units, F = l2 - l1).
// The wave works in a workspace (LDS when it fits the launch's per-wave size, else
// the pair's slice of scratch):
//  * match masks: X[k][t], the positions of N's word k holding H[t], for every t, and
//    R[k][j], the same for the reversed needle against H[l2 - 1 - j], j < l1.  8-bit
//    needles of at most two words read them from per-symbol tables (R bit-reversed);
//    the others take one ballot per word per haystack unit;
//  * the LCS with N of every candidate window, from the bit-parallel LCS recurrence
//    (Hyyro 2004), one window per lane: the full windows [p, p + l1), p < F, each
//    l1 steps; all prefixes [0, i) in one pass of one lane over H (the recurrence's
//    state after i units is the prefix's LCS) and all suffixes [i, l2) in one pass of
//    another lane over the reversed strings;
//  * the reference's choice among them.
//
// Full windows: the reference bisects [0, F - 1] level by level, visiting each range
// at both ends and splitting it (at the floor of its middle) while the bound
// min(d_f, d_sc) - improvement on the distances inside can beat the running cutoff.
// Window distances are even and change by at most 2 per unit of shift, so the bound
// never exceeds a distance inside: every window below the running cutoff lies in a
// range that is split, the search reaches the smallest distance g over all full
// windows (when g beats the entry cutoff), and until it does every range around a
// window of distance g is split.  The first window of distance g it visits is
// therefore the one that first becomes a range end in the complete halving tree of
// [0, F - 1] -- the smallest depth, then the smallest position (a level's ends are
// visited left to right).  The kernel takes g from a reduction of the windows' LCS
// and that window from a depth computation over the windows of distance g.
//
// Prefixes, then suffixes: a window replaces the result when its ratio (0 below the
// running cutoff) is strictly larger, so a phase keeps the first window of the
// largest ratio among those whose boundary unit occurs in N and whose ratio reaches
// the cutoff -- one max / first-index reduction over the lanes.
// Equal lengths: the same search with the strings' roles exchanged, kept when
// strictly better (its windows lie in the first string).
#include "rfa_block_impl.hpp"
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"

namespace rfa {

// CachedRatio(needle).similarity(window, cutoff) (fuzz_impl.hpp:42-47) from the LCS
__device__ __forceinline__ double pr_ratio(uint64_t l1, uint64_t l2, uint64_t lcs, double cutoff)
{
    ScoreSpec t{};
    t.metric = RFA_INDEL;
    const IntPair q{&t, l1, l2, lcs};
    return int_api_norm_similarity(q, cutoff / 100) * 100;
}

struct PrResult {
    double score;
    uint32_t ds, de; // the window of the haystack
};

// The needle's match masks over the haystack: xtab[k * l2 + t] and, for the reversed
// needle against the haystack read backwards, xrev[k * l1 + j].
template <uint32_t WM, typename CharT, bool LDS>
__device__ __forceinline__ void pr_masks(const CharT* __restrict__ nd, uint32_t l1, const CharT* __restrict__ hy,
                                         uint32_t l2, uint32_t W, bool tables, uint64_t* xtab, uint64_t* xrev,
                                         uint64_t* sym)
{
    const uint32_t lane = threadIdx.x & 63;
    if constexpr (LDS && WM <= 4 && sizeof(CharT) == 1) {
        if (tables) { // W <= 2
            // per-symbol masks of the needle: sym[c * W + k]
            for (uint32_t i = lane; i < 256 * W; i += 64) sym[i] = 0;
            __threadfence_block();
            for (uint32_t k = 0; k < W; ++k) {
                const uint32_t i = 64 * k + lane;
                if (i < l1) atomicOr((unsigned long long*)&sym[(uint32_t)nd[i] * W + k], 1ull << lane);
            }
            __threadfence_block();
            // 512 haystack units per round, their loads issued together
            for (uint32_t t0 = 0; t0 < l2; t0 += 512) {
                uint32_t c[8];
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q) {
                    const uint32_t t = t0 + 64 * q + lane;
                    c[q] = t < l2 ? (uint32_t)hy[t] : 0u;
                }
#pragma unroll
                for (uint32_t q = 0; q < 8; ++q) {
                    const uint32_t t = t0 + 64 * q + lane;
                    if (t < l2)
                        for (uint32_t k = 0; k < W; ++k) xtab[(uint64_t)k * l2 + t] = sym[c[q] * W + k];
                }
            }
            // the reversed needle's masks: the needle's, bit-reversed within l1 bits
            for (uint32_t j = lane; j < l1; j += 64) {
                const uint32_t c = (uint32_t)hy[l2 - 1 - j];
                if (W == 1) {
                    xrev[j] = __builtin_bitreverse64(sym[c]) >> (64 - l1);
                }
                else {
                    const uint64_t hr = __builtin_bitreverse64(sym[2 * c]), lr = __builtin_bitreverse64(sym[2 * c + 1]);
                    const uint32_t sh = 128 - l1; // 0..63
                    xrev[j] = sh ? (lr >> sh) | (hr << (64 - sh)) : lr;
                    xrev[l1 + j] = hr >> sh;
                }
            }
            __threadfence_block();
            return;
        }
    }
    uint32_t nu[WM], nr[WM];
    bool nv[WM];
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) {
        const uint32_t i = 64 * k + lane;
        nv[k] = k < W && i < l1;
        nu[k] = nv[k] ? (uint32_t)nd[i] : 0u;
        nr[k] = nv[k] ? (uint32_t)nd[l1 - 1 - i] : 0u;
    }
    for (int dir = 0; dir < 2; ++dir) {
        const uint32_t n = dir == 0 ? l2 : l1;
        uint64_t* out = dir == 0 ? xtab : xrev;
        for (uint32_t t0 = 0; t0 < n; t0 += 64) {
            const uint32_t cnt = min(64u, n - t0);
            const uint32_t h = lane < cnt ? (uint32_t)hy[dir == 0 ? t0 + lane : l2 - 1 - (t0 + lane)] : 0u;
            uint64_t acc[WM];
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) acc[k] = 0;
            for (uint32_t i = 0; i < cnt; ++i) {
                const uint32_t c = readlane(h, i);
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k) {
                    if (k < W) {
                        const uint64_t m = ballot(nv[k] && (dir == 0 ? nu[k] : nr[k]) == c);
                        acc[k] = lane == i ? m : acc[k];
                    }
                }
            }
            if (lane < cnt) {
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k)
                    if (k < W) out[(uint64_t)k * n + t0 + lane] = acc[k];
            }
        }
    }
    __threadfence_block();
}

// l1 steps of the LCS recurrence over p[k * stride + j] (every lane steps; the lanes
// past their windows run on in-bounds words and are not read); a recorder lane stores
// the LCS after each of its first len steps at lcs[recbase + recdir * j].  The next
// step's words are loaded before the current step's arithmetic.
template <uint32_t WM, bool REC>
__device__ __forceinline__ void pr_lcs_run(const uint64_t* p, uint32_t stride, uint32_t len, uint32_t l1, uint32_t W,
                                           const uint64_t (&lastm)[WM], uint64_t (&S)[WM], uint32_t* lcs,
                                           int recbase, int recdir, bool isrec)
{
    uint64_t Xn[WM];
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) Xn[k] = k < W ? p[(uint64_t)k * stride] : 0ull;
    for (uint32_t j = 0; j < l1; ++j) {
        uint64_t X[WM];
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k) X[k] = Xn[k];
        const uint32_t jn = j + 1 < l1 ? j + 1 : j;
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k)
            if (k < W) Xn[k] = p[(uint64_t)k * stride + jn];
        unsigned long long carry = 0;
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k) {
            if (k < W) {
                const uint64_t u = S[k] & X[k];
                unsigned long long co;
                const uint64_t x = __builtin_addcll(S[k], u, carry, &co);
                carry = co;
                S[k] = x | (S[k] & ~u);
            }
        }
        if constexpr (REC) {
            if (isrec && j < len) {
                uint32_t v = 0;
#pragma unroll
                for (uint32_t k = 0; k < WM; ++k)
                    if (k < W) v += (uint32_t)__popcll(~S[k] & lastm[k]);
                lcs[recbase + recdir * (int)j] = v;
            }
        }
    }
}

// One search (fuzz_detail::partial_ratio_impl, fuzz_impl.hpp:68-179) of needle nd
// (l1 <= WM * 64 units) in haystack hy (l2 >= l1 units) in workspace ws
// (pralign_ws_bytes(l1, l2, tables) bytes).  The result is uniform across the wave.
template <uint32_t WM, typename CharT, bool LDS>
__device__ __forceinline__ PrResult pr_search(const CharT* __restrict__ nd, uint32_t l1, const CharT* __restrict__ hy,
                                              uint32_t l2, double cutoff, uint64_t* ws, bool tables)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = WM <= 2 ? WM : (l1 + 63) / 64; // the classes of 1 and 2 words are exact
    const uint32_t F = l2 - l1;
    uint64_t* xtab = ws;
    uint64_t* xrev = ws + (uint64_t)W * l2;
    uint64_t* rest = xrev + (uint64_t)W * l1;
    uint32_t* lcs = (uint32_t*)rest; // F full windows, l1 - 1 prefixes, l1 suffixes

    pr_masks<WM, CharT, LDS>(nd, l1, hy, l2, W, tables, xtab, xrev, rest);

    // ---- LCS of every candidate window ----
    uint64_t lastm[WM]; // the needle's bits of each word
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) {
        const uint32_t lo = 64 * k;
        lastm[k] = k >= W ? 0ull : (l1 - lo >= 64 ? ~0ull : ((1ull << (l1 - lo)) - 1));
    }
    uint32_t gmax = 0; // the lane's largest LCS of a full window
    for (uint32_t r0 = 0; r0 < F + 2; r0 += 64) {
        const uint32_t r = r0 + lane;
        const uint64_t* p = xtab;
        uint32_t stride = l2, len = 0;
        int recbase = 0, recdir = 0;
        bool isrec = false;
        if (r < F) {
            p = xtab + r;
            len = l1;
        }
        else if (r == F) { // prefixes [0, j + 1)
            len = l1 - 1;
            isrec = true;
            recbase = (int)F;
            recdir = 1;
        }
        else if (r == F + 1) { // suffixes [l2 - 1 - j, l2)
            p = xrev;
            stride = l1;
            len = l1;
            isrec = true;
            recbase = (int)(F + 2 * l1 - 2);
            recdir = -1;
        }
        uint64_t S[WM];
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k) S[k] = ~0ull;
        if (r0 + 64 > F)
            pr_lcs_run<WM, true>(p, stride, len, l1, W, lastm, S, lcs, recbase, recdir, isrec);
        else
            pr_lcs_run<WM, false>(p, stride, len, l1, W, lastm, S, lcs, recbase, recdir, isrec);
        if (r < F) {
            uint32_t v = 0;
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k)
                if (k < W) v += (uint32_t)__popcll(~S[k] & lastm[k]);
            lcs[r] = v;
            gmax = max(gmax, v);
        }
    }
    __threadfence_block();

    // ---- full windows: the window the reference's bisection settles on ----
    PrResult res{0.0, 0u, l1};
    if (F > 0) {
        const uint32_t maximum = 2 * l1;
        // normsim_to_normdist is at most 1: the cutoff distance is at most maximum
        const uint32_t cut = (uint32_t)ceil((double)maximum * normsim_to_normdist(cutoff / 100));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) gmax = max(gmax, (uint32_t)__shfl_xor((int)gmax, off, 64));
        const uint32_t G = uniform(gmax), g = maximum - 2 * G;
        uint64_t best = ~0ull; // no window beat the cutoff
        if (g < cut) {
            // among the windows of distance g: smallest depth in the halving tree, then position
            uint32_t bdep = ~0u, bpos = 0;
            for (uint32_t r0 = 0; r0 < F; r0 += 64) {
                const uint32_t p = r0 + lane;
                const bool cand = p < F && lcs[p] == G;
                if (ballot(cand) == 0) continue;
                uint32_t dep = ~0u;
                if (cand) {
                    dep = 0;
                    if (p != 0 && p != F - 1) {
                        uint32_t lo = 0, hi = F - 1; // lo < p < hi throughout
                        for (;;) {
                            const uint32_t mid = lo + ((hi - lo) >> 1);
                            ++dep;
                            if (p == mid) break;
                            if (p < mid)
                                hi = mid;
                            else
                                lo = mid;
                        }
                    }
                }
                uint32_t md = dep;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) md = min(md, (uint32_t)__shfl_xor((int)md, off, 64));
                md = uniform(md);
                if (md < bdep) { // later chunks hold later positions: ties keep the earlier
                    bdep = md;
                    bpos = r0 + (uint32_t)__builtin_ctzll(ballot(dep == md));
                }
            }
            res.ds = bpos;
            res.de = bpos + l1;
            if (g == 0) {
                res.score = 100.0;
                return res;
            }
            best = g;
        }
        double sv = 1.0 - ((double)best / (double)maximum);
        sv *= 100;
        if (sv >= cutoff) {
            res.score = sv;
            cutoff = sv;
        }
    }

    // ---- prefixes, then suffixes: first window of the largest ratio ----
    for (int phase = 0; phase < 2; ++phase) {
        const uint32_t n = phase == 0 ? l1 - 1 : l1; // windows of the phase
        double best = -1.0;
        uint32_t bidx = ~0u;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            if (i < n) {
                // prefix [0, i + 1) (boundary unit H[i]); suffix [F + i, l2) (boundary H[F + i])
                const uint32_t t = phase == 0 ? i : F + i;
                const uint32_t wl = phase == 0 ? i + 1 : l2 - (F + i);
                bool in = false;
                for (uint32_t k = 0; k < W; ++k) in = in || xtab[(uint64_t)k * l2 + t] != 0;
                if (in) {
                    const uint32_t v = lcs[(phase == 0 ? F : F + l1 - 1) + i];
                    const double ls = pr_ratio(l1, wl, v, cutoff);
                    if (ls > best) {
                        best = ls;
                        bidx = i;
                    }
                }
            }
        }
        // the wave's largest ratio, the smallest index among its lanes
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const uint32_t oi = (uint32_t)__shfl_xor((int)bidx, off, 64);
            if (ob > best || (ob == best && oi < bidx)) {
                best = ob;
                bidx = oi;
            }
        }
        if (bidx != ~0u && best > res.score) {
            res.score = best;
            if (phase == 0) {
                res.ds = 0;
                res.de = bidx + 1;
            }
            else {
                res.ds = F + bidx;
                res.de = l2;
            }
            cutoff = best;
            if (res.score == 100.0) return res;
        }
    }
    return res;
}

// partial_ratio_alignment of one pair (fuzz_impl.hpp:196-230) in workspace ws
template <uint32_t WM, typename CharT, bool LDS>
__device__ __forceinline__ void pr_pair(const PrAlignArgs& a, uint32_t pair, const CharT* A, uint32_t la,
                                        const CharT* B, uint32_t lb, uint64_t* ws, bool tables)
{
    const uint32_t lane = threadIdx.x & 63;
    const bool swap = la > lb; // partial_ratio_alignment(s2, s1) with the fields exchanged
    const CharT* nd = swap ? B : A;
    const CharT* hy = swap ? A : B;
    const uint32_t l1 = swap ? lb : la, l2 = swap ? la : lb;
    uint32_t ss = 0, se = l1, dst = 0, den = l1;
    double score = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        double c = a.cutoff;
        if (pass == 1) {
            if (score == 100.0 || l1 != l2) break;
            c = a.cutoff < score ? score : a.cutoff; // std::max(score_cutoff, score)
            __threadfence_block();
        }
        // the second pass searches the first string for the second (equal lengths)
        const PrResult r = pr_search<WM, CharT, LDS>(pass ? hy : nd, l1, pass ? nd : hy, l2, c, ws, tables);
        if (pass == 0) {
            score = r.score;
            dst = r.ds;
            den = r.de;
        }
        else if (r.score > score) {
            score = r.score;
            ss = r.ds;
            se = r.de;
            dst = 0;
            den = l2;
        }
    }
    // lanes 0-4 write the pair's rfa_score_alignment
    if (lane < 5) {
        uint64_t v = (uint64_t)__double_as_longlong(score);
        if (lane == 1) v = swap ? dst : ss;
        if (lane == 2) v = swap ? den : se;
        if (lane == 3) v = swap ? ss : dst;
        if (lane == 4) v = swap ? se : den;
        a.out[5ull * pair + lane] = v;
    }
}

template <uint32_t WM, typename CharT>
__global__ __launch_bounds__(64 * PRALIGN_WAVES) void k_pralign(PrAlignArgs a)
{
    extern __shared__ uint64_t pr_lds[];
    // the pair and its lengths are wave-uniform: keep them (and the loops they bound) scalar
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t item = blockIdx.x * PRALIGN_WAVES + wave;
    if (item >= a.nitems) return;
    const uint32_t pair = uniform(a.items[item]);
    const uint64_t ao = a.aoff[pair], bo = a.boff[pair];
    const CharT* A = (const CharT*)a.achars + ao;
    const CharT* B = (const CharT*)a.bchars + bo;
    const uint32_t la = uniform((uint32_t)(a.aoff[pair + 1] - ao)), lb = uniform((uint32_t)(a.boff[pair + 1] - bo));
    const uint32_t l1 = min(la, lb), l2 = max(la, lb);
    const bool tables = pralign_symbol_tables((int)sizeof(CharT), l1);
    if constexpr (WM <= 4) { // wider needles never fit the LDS cap
        if (pralign_ws_bytes(l1, l2, tables) <= a.lds_bytes) {
            pr_pair<WM, CharT, true>(a, pair, A, la, B, lb, pr_lds + (uint64_t)wave * (a.lds_bytes / 8), tables);
            return;
        }
    }
    pr_pair<WM, CharT, false>(a, pair, A, la, B, lb, (uint64_t*)(a.scratch + a.scratch_off[item]), false);
}

// ---- needles of 1025..4096 units (k_pbig) ---------------------------------------
//
// One wave per pair, the needle's words across the lanes (word w in lane w), its PM rows
// and its reversal's in the wave's workspace (build_pattern); every haystack unit's row
// index is looked up once.  A window's LCS runs the bit-parallel recurrence over the
// words with the additions carried across lanes by lookahead on two ballots, the match
// words of eight steps loaded ahead of their use.  The full windows follow the
// reference's bisection itself (fuzz_impl.hpp:68-146), one window at a time with its
// distance kept for revisits; the prefixes and the suffixes are recorded in one pass
// each (the reversed needle against the haystack read backwards) and chosen like the
// one-word-per-lane kernel above.

template <typename CharT, bool REC>
__device__ uint32_t pb_lcs(const Pattern<CharT, false>& P, uint32_t l1, const uint32_t* __restrict__ rowseq,
                           uint32_t n, uint32_t* __restrict__ rec)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = P.W;
    const bool act = lane < W;
    const uint64_t valid = !act ? 0ull : (l1 - 64ull * lane >= 64 ? ~0ull : ((1ull << (l1 - 64ull * lane)) - 1));
    const uint64_t* __restrict__ pm = P.pm;
    const uint32_t Wp = P.Wp;
    uint64_t S = ~0ull;
    for (uint32_t j0 = 0; j0 < n; j0 += 64) {
        const uint32_t rr = j0 + lane < n ? rowseq[j0 + lane] : 0u;
        const uint32_t steps = min(64u, n - j0);
        uint64_t Xc[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            const uint32_t r = q < steps ? readlane(rr, q) : 0u;
            Xc[q] = (act && q < steps) ? pm[(size_t)r * Wp + lane] : 0ull;
        }
        for (uint32_t g0 = 0; g0 < steps; g0 += 8) {
            uint64_t Xn[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                const uint32_t s2 = g0 + 8 + q;
                const uint32_t r = s2 < steps ? readlane(rr, s2) : 0u;
                Xn[q] = (act && s2 < steps) ? pm[(size_t)r * Wp + lane] : 0ull;
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                if (g0 + q < steps) {
                    const uint64_t u = S & Xc[q];
                    uint64_t x = S + u;
                    const bool g = act && (x < S);
                    const bool pr = act && !g && (x == ~0ull);
                    const uint64_t G = ballot(g), Pp = ballot(pr);
                    const uint64_t sum = (G | Pp) + G;
                    x += ((sum ^ Pp) >> lane) & 1u;
                    if (act) S = x | (S - u);
                    if constexpr (REC) {
                        uint32_t v = (uint32_t)__popcll(~S & valid);
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
                        if (lane == 0) rec[j0 + g0 + q] = v;
                    }
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) Xc[q] = Xn[q];
        }
    }
    uint32_t v = (uint32_t)__popcll(~S & valid);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
    return v;
}

// fuzz_detail::partial_ratio_impl (fuzz_impl.hpp:68-179) of needle nd in haystack hy
template <typename CharT>
__device__ PrResult pb_search(const CharT* __restrict__ nd, uint32_t l1, const CharT* __restrict__ hy, uint32_t l2,
                              double cutoff, uint8_t* ws)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t o[10];
    pbig_layout((int)sizeof(CharT), l1, l2, o);
    CharT* rev = (CharT*)(ws + o[2]);
    for (uint32_t i = lane; i < l1; i += 64) rev[i] = nd[l1 - 1 - i];
    arena_sync<false>();
    const Pattern<CharT, false> P = build_pattern<CharT, false>(nd, l1, ws + o[0], pattern_hash_bits<CharT>(l1));
    const Pattern<CharT, false> R = build_pattern<CharT, false>(rev, l1, ws + o[1], pattern_hash_bits<CharT>(l1));
    uint32_t* rows = (uint32_t*)(ws + o[3]);
    uint32_t* rrows = (uint32_t*)(ws + o[4]);
    for (uint32_t t = lane; t < l2; t += 64) rows[t] = P.map.find((uint32_t)hy[t]);
    for (uint32_t j = lane; j < l1; j += 64) rrows[j] = R.map.find((uint32_t)hy[l2 - 1 - j]);
    arena_sync<false>();

    PrResult res{0.0, 0u, l1};
    const uint32_t F = l2 - l1;
    if (F > 0) {
        uint32_t* scores = (uint32_t*)(ws + o[5]);
        uint2* cur = (uint2*)(ws + o[6]);
        uint2* nxt = cur + (2 * F + 2);
        constexpr uint32_t UNSET = 0xffffffffu;
        for (uint32_t i = lane; i < F; i += 64) scores[i] = UNSET;
        if (lane == 0) cur[0] = make_uint2(0u, F - 1);
        arena_sync<false>();
        const uint64_t maximum = 2ull * l1;
        uint64_t cutoff_dist = (uint64_t)ceil((double)maximum * normsim_to_normdist(cutoff / 100));
        uint64_t best = ~0ull;
        uint32_t ncur = 1, ds = 0, de = l1;
        bool perfect = false;
        // distance of window p: evaluated once, kept for revisits (the last one in registers:
        // adjacent ranges of a level share an end; earlier levels' are read after a barrier)
        uint32_t last_p = UNSET, last_d = 0;
        auto dist_of = [&](uint32_t p, bool& fresh) -> uint64_t {
            if (p == last_p) {
                fresh = false;
                return last_d;
            }
            const uint32_t known = uniform(scores[p]);
            fresh = known == UNSET;
            if (!fresh) return known;
            const uint32_t d = (uint32_t)(maximum - 2ull * pb_lcs<CharT, false>(P, l1, rows + p, l1, nullptr));
            if (lane == 0) scores[p] = d;
            last_p = p;
            last_d = d;
            return d;
        };
        while (ncur && !perfect) {
            uint32_t nnext = 0;
            for (uint32_t k = 0; k < ncur && !perfect; ++k) {
                const uint32_t f = uniform(cur[k].x), sc = uniform(cur[k].y);
                bool fresh;
                const uint64_t df = dist_of(f, fresh);
                if (fresh && df < cutoff_dist) {
                    cutoff_dist = best = df;
                    ds = f;
                    de = f + l1;
                    if (best == 0) perfect = true;
                }
                if (perfect) break;
                const uint64_t dsc = dist_of(sc, fresh);
                if (fresh && dsc < cutoff_dist) {
                    cutoff_dist = best = dsc;
                    ds = sc;
                    de = sc + l1;
                    if (best == 0) perfect = true;
                }
                if (perfect) break;
                const uint32_t cell_diff = sc - f;
                if (cell_diff == 1) continue;
                const uint64_t known = df > dsc ? df - dsc : dsc - df;
                const uint64_t improve = ((uint64_t)cell_diff - known / 2) / 2 * 2;
                const int64_t min_score = (int64_t)(df < dsc ? df : dsc) - (int64_t)improve;
                if (min_score < (int64_t)cutoff_dist) {
                    const uint32_t center = cell_diff / 2;
                    if (lane == 0) {
                        nxt[nnext] = make_uint2(f, f + center);
                        nxt[nnext + 1] = make_uint2(f + center, sc);
                    }
                    nnext += 2;
                }
            }
            arena_sync<false>(); // the next level's list and the distances
            uint2* t = cur;
            cur = nxt;
            nxt = t;
            ncur = perfect ? 0 : nnext;
        }
        res.ds = ds;
        res.de = de;
        if (perfect) {
            res.score = 100.0;
            return res;
        }
        double sv = 1.0 - ((double)best / (double)maximum);
        sv *= 100;
        if (sv >= cutoff) {
            res.score = sv;
            cutoff = sv;
        }
    }

    // ---- prefixes [0, i + 1) and suffixes [F + i, l2): recorded, then chosen ----
    uint32_t* rec_pre = (uint32_t*)(ws + o[7]);
    uint32_t* rec_suf = (uint32_t*)(ws + o[8]);
    if (l1 > 1) pb_lcs<CharT, true>(P, l1, rows, l1 - 1, rec_pre);
    pb_lcs<CharT, true>(R, l1, rrows, l1, rec_suf); // after step j: the suffix [l2 - 1 - j, l2)
    arena_sync<false>();
    for (int phase = 0; phase < 2; ++phase) {
        const uint32_t n = phase == 0 ? l1 - 1 : l1;
        double bestv = -1.0;
        uint32_t bidx = ~0u;
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + lane;
            if (i < n) {
                const uint32_t t = phase == 0 ? i : F + i;
                const uint32_t wl = phase == 0 ? i + 1 : l2 - (F + i);
                if (rows[t] != 0) { // the boundary unit occurs in the needle
                    const uint32_t v = phase == 0 ? rec_pre[i] : rec_suf[l1 - 1 - i];
                    const double ls = pr_ratio(l1, wl, v, cutoff);
                    if (ls > bestv) {
                        bestv = ls;
                        bidx = i;
                    }
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(bestv, off, 64);
            const uint32_t oi = (uint32_t)__shfl_xor((int)bidx, off, 64);
            if (ob > bestv || (ob == bestv && oi < bidx)) {
                bestv = ob;
                bidx = oi;
            }
        }
        if (bidx != ~0u && bestv > res.score) {
            res.score = bestv;
            if (phase == 0) {
                res.ds = 0;
                res.de = bidx + 1;
            }
            else {
                res.ds = F + bidx;
                res.de = l2;
            }
            cutoff = bestv;
            if (res.score == 100.0) return res;
        }
    }
    return res;
}

template <typename CharT>
__global__ __launch_bounds__(64) void k_pbig(PbigArgs a)
{
    const uint32_t lane = threadIdx.x & 63;
    uint8_t* ws = a.ws + (uint64_t)blockIdx.x * a.ws_stride;
    for (uint64_t it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        uint32_t ia, ib;
        if (a.cdist) {
            ia = a.items[it / a.nb];
            ib = (uint32_t)(it % a.nb);
        }
        else {
            ia = ib = a.items[it];
        }
        ia = uniform(ia);
        ib = uniform(ib);
        const uint64_t ao = a.aoff[ia], bo = a.boff[ib];
        const CharT* A = (const CharT*)a.achars + ao;
        const CharT* B = (const CharT*)a.bchars + bo;
        const uint32_t la = uniform((uint32_t)(a.aoff[ia + 1] - ao)), lb = uniform((uint32_t)(a.boff[ib + 1] - bo));
        const uint32_t l1 = min(la, lb), l2 = max(la, lb);
        if (a.cdist && l1 <= PRALIGN_MAX) continue; // the long partial kernel's pairs
        const double cut = a.cdist && a.pcut ? a.pcut[(uint64_t)ia * a.out_stride + ib] : a.cutoff;
        const bool swap = la > lb;
        const CharT* nd = swap ? B : A;
        const CharT* hy = swap ? A : B;
        uint32_t ss = 0, se = l1, dst = 0, den = l1;
        double score = 0.0;
        if (cut <= 100) {
            const PrResult r = pb_search<CharT>(nd, l1, hy, l2, cut, ws);
            score = r.score;
            dst = r.ds;
            den = r.de;
            if (score != 100.0 && l1 == l2) {
                const double c2 = cut < score ? score : cut;
                arena_sync<false>();
                const PrResult r2 = pb_search<CharT>(hy, l2, nd, l1, c2, ws);
                if (r2.score > score) {
                    score = r2.score;
                    ss = r2.ds;
                    se = r2.de;
                    dst = 0;
                    den = l2;
                }
            }
        }
        if (a.cdist) {
            if (lane == 0) a.score_out[(uint64_t)ia * a.out_stride + ib] = score;
        }
        else if (lane < 5) {
            uint64_t v = (uint64_t)__double_as_longlong(score);
            if (lane == 1) v = swap ? dst : ss;
            if (lane == 2) v = swap ? den : se;
            if (lane == 3) v = swap ? ss : dst;
            if (lane == 4) v = swap ? se : den;
            a.al_out[5ull * ia + lane] = v;
        }
        arena_sync<false>(); // the workspace is reused by the next item
    }
}

hipError_t launch_pbig(int char_width, const PbigArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (a.nitems == 0) return hipSuccess;
    if (char_width == 1)
        hipLaunchKernelGGL((k_pbig<uint8_t>), dim3(nblocks), dim3(64), 0, s, a);
    else if (char_width == 2)
        hipLaunchKernelGGL((k_pbig<uint16_t>), dim3(nblocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_pbig<uint32_t>), dim3(nblocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <uint32_t WM>
static void launch_w(int char_width, const PrAlignArgs& a, uint32_t blocks, hipStream_t s)
{
    const size_t lds = (size_t)PRALIGN_WAVES * a.lds_bytes;
    if (char_width == 1)
        hipLaunchKernelGGL((k_pralign<WM, uint8_t>), dim3(blocks), dim3(64 * PRALIGN_WAVES), lds, s, a);
    else if (char_width == 2)
        hipLaunchKernelGGL((k_pralign<WM, uint16_t>), dim3(blocks), dim3(64 * PRALIGN_WAVES), lds, s, a);
    else
        hipLaunchKernelGGL((k_pralign<WM, uint32_t>), dim3(blocks), dim3(64 * PRALIGN_WAVES), lds, s, a);
}

hipError_t launch_pralign(int char_width, uint32_t words, const PrAlignArgs& a, hipStream_t s)
{
    if (a.nitems == 0) return hipSuccess;
    if (a.lds_bytes > PRALIGN_LDS_CAP || a.lds_bytes % 8) return hipErrorInvalidValue;
    const uint32_t blocks = (a.nitems + PRALIGN_WAVES - 1) / PRALIGN_WAVES;
    if (words <= 1)
        launch_w<1>(char_width, a, blocks, s);
    else if (words <= 2)
        launch_w<2>(char_width, a, blocks, s);
    else if (words <= 4)
        launch_w<4>(char_width, a, blocks, s);
    else if (words <= 8)
        launch_w<8>(char_width, a, blocks, s);
    else
        launch_w<16>(char_width, a, blocks, s);
    return hipGetLastError();
}

} // namespace rfa
