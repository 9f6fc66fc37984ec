// rfa_pralign.hip -- fuzz::partial_ratio_alignment of pairs (fuzz.hpp:129-135,
// fuzz_impl.hpp:68-230): the best partial ratio and the windows it was found at.
//
// One wave per pair.  The needle N (the shorter string, l1 units; the first string
// when the lengths are equal) sits one unit per lane per 64-unit word.  For every
// haystack position t one ballot per word gives the positions of N holding H[t]
// (X[k][t]); the table goes to the pair's scratch, word-major, so both its writes
// and the lanes' later reads are coalesced.  Every candidate window of H -- the
// full windows [p, p + l1) for p < l2 - l1, the prefixes [0, i) for 0 < i < l1 and
// the suffixes [i, l2) for l2 - l1 <= i < l2 -- gets its LCS with N from the
// bit-parallel LCS recurrence (Hyyro 2004), one window per lane.
//
// Then the reference's choice among them:
//  * full windows: its bisection of [0, l2 - l1 - 1] (the bound on the distances
//    between two evaluated windows, the running cutoff), replayed by lane 0 on the
//    stored LCS values with window lists in scratch: the result is the first window
//    visited whose distance beats the cutoff at the time;
//  * prefixes, then suffixes: a window replaces the result when its ratio (0 below
//    the running cutoff) is strictly larger, so the phase keeps the first window of
//    the largest ratio among those whose boundary unit occurs in N and whose ratio
//    reaches the cutoff -- one max / first-index reduction over the lanes;
//  * equal lengths: the same search with the strings' roles exchanged, kept when
//    strictly better (its windows lie in the first string).
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

// One search (fuzz_detail::partial_ratio_impl, fuzz_impl.hpp:68-179) of needle nd
// (l1 <= WM * 64 units) in haystack hy (l2 >= l1 units).  xtab: WM * l2 words;
// lcs: l2 - l1 + 2 * l1 entries; lists: 2 * (2 * (l2 - l1) + 2) windows.  The result
// is uniform across the wave.
template <uint32_t WM, typename CharT>
__device__ PrResult pr_search(const CharT* __restrict__ nd, uint32_t l1, const CharT* __restrict__ hy, uint32_t l2,
                              double cutoff, uint64_t* __restrict__ xtab, uint32_t* __restrict__ lcs,
                              uint2* __restrict__ lists)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = (l1 + 63) / 64;
    // ---- the needle's match masks over the haystack ----
    uint32_t nu[WM];
    bool nv[WM];
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) {
        nv[k] = k < W && 64 * k + lane < l1;
        nu[k] = nv[k] ? (uint32_t)nd[64 * k + lane] : 0u;
    }
    for (uint32_t t0 = 0; t0 < l2; t0 += 64) {
        const uint32_t cnt = min(64u, l2 - t0);
        const uint32_t h = lane < cnt ? (uint32_t)hy[t0 + lane] : 0u;
        uint64_t acc[WM];
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k) acc[k] = 0;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t c = readlane(h, i);
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) {
                const uint64_t m = ballot(nv[k] && nu[k] == c);
                acc[k] = lane == i ? m : acc[k];
            }
        }
        if (lane < cnt) {
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k)
                if (k < W) xtab[(uint64_t)k * l2 + t0 + lane] = acc[k];
        }
    }
    __threadfence_block();

    // ---- LCS of every candidate window, one per lane ----
    const uint32_t F = l2 - l1; // full windows [p, p + l1), p < F
    const uint32_t nwin = F + 2 * l1 - 1;
    uint64_t lastm[WM]; // the needle's bits of each word
#pragma unroll
    for (uint32_t k = 0; k < WM; ++k) {
        const uint32_t lo = 64 * k;
        lastm[k] = k >= W ? 0ull : (l1 - lo >= 64 ? ~0ull : ((1ull << (l1 - lo)) - 1));
    }
    for (uint32_t r0 = 0; r0 < nwin; r0 += 64) {
        const uint32_t r = r0 + lane;
        uint32_t s = 0, e = 0;
        if (r < F) {
            s = r;
            e = r + l1;
        }
        else if (r < F + l1 - 1) {
            e = r - F + 1; // prefix [0, e)
        }
        else if (r < nwin) {
            s = (r - F - (l1 - 1)) + F; // suffix [s, l2)
            e = l2;
        }
        uint32_t len = e - s, lmax = len;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) lmax = max(lmax, (uint32_t)__shfl_xor(lmax, off, 64));
        lmax = uniform(lmax);
        uint64_t S[WM];
#pragma unroll
        for (uint32_t k = 0; k < WM; ++k) S[k] = ~0ull;
        for (uint32_t j = 0; j < lmax; ++j) {
            const bool act = j < len;
            const uint32_t t = act ? s + j : 0u;
            uint64_t carry = 0;
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) {
                if (k < W) {
                    const uint64_t X = xtab[(uint64_t)k * l2 + t];
                    const uint64_t u = S[k] & X;
                    const uint64_t x = S[k] + u + carry;
                    carry = (x < S[k] || (carry && x == S[k])) ? 1ull : 0ull;
                    const uint64_t ns = x | (S[k] - u);
                    S[k] = act ? ns : S[k];
                }
            }
        }
        if (r < nwin) {
            uint32_t v = 0;
#pragma unroll
            for (uint32_t k = 0; k < WM; ++k) v += (uint32_t)__popcll(~S[k] & lastm[k]);
            lcs[r] = v;
        }
    }
    __threadfence_block();

    // ---- full windows: the reference's bisection (lane 0) ----
    PrResult res{0.0, 0u, l1};
    if (l2 > l1) {
        double score = 0.0;
        bool perfect = false;
        uint32_t ds = 0, de = l1;
        if (lane == 0) {
            const uint64_t maximum = 2ull * l1;
            uint64_t cutoff_dist = (uint64_t)ceil((double)maximum * normsim_to_normdist(cutoff / 100));
            uint64_t best = ~0ull;
            constexpr uint32_t SEEN = 0x80000000u;
            auto dist_at = [&](uint32_t p, bool& fresh) -> uint64_t {
                const uint32_t v = lcs[p];
                fresh = (v & SEEN) == 0;
                if (fresh) lcs[p] = v | SEEN;
                return maximum - 2ull * (v & ~SEEN);
            };
            uint2* cur = lists;
            uint2* nxt = lists + (2 * F + 2);
            uint32_t ncur = 1, nnext = 0;
            cur[0] = make_uint2(0u, F - 1);
            while (ncur && !perfect) {
                for (uint32_t k = 0; k < ncur && !perfect; ++k) {
                    const uint32_t f = cur[k].x, sc = cur[k].y;
                    bool fresh;
                    const uint64_t df = dist_at(f, fresh);
                    if (fresh && df < cutoff_dist) {
                        cutoff_dist = best = df;
                        ds = f;
                        de = f + l1;
                        if (best == 0) perfect = true;
                    }
                    if (perfect) break;
                    const uint64_t dsc = dist_at(sc, fresh);
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
                    const uint64_t improve = (cell_diff - known / 2) / 2 * 2;
                    const int64_t min_score = (int64_t)(df < dsc ? df : dsc) - (int64_t)improve;
                    if (min_score < (int64_t)cutoff_dist) {
                        const uint32_t center = cell_diff / 2;
                        nxt[nnext++] = make_uint2(f, f + center);
                        nxt[nnext++] = make_uint2(f + center, sc);
                    }
                }
                uint2* tsw = cur;
                cur = nxt;
                nxt = tsw;
                ncur = nnext;
                nnext = 0;
            }
            if (perfect) {
                score = 100.0;
            }
            else {
                double sv = 1.0 - ((double)best / (double)maximum);
                sv *= 100;
                if (sv >= cutoff) score = sv;
            }
        }
        // lane 0's values to the wave (the SEEN flags stay in scratch only)
        res.score = __shfl(score, 0, 64);
        res.ds = (uint32_t)__shfl((int)ds, 0, 64);
        res.de = (uint32_t)__shfl((int)de, 0, 64);
        perfect = __shfl((int)perfect, 0, 64) != 0;
        if (perfect) return res;
        if (res.score >= cutoff) cutoff = res.score; // score_cutoff = res.score when it was set
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
                    const uint32_t v = lcs[(phase == 0 ? F : F + l1 - 1) + i] & 0x7fffffffu;
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

template <uint32_t WM, typename CharT>
__global__ __launch_bounds__(256) void k_pralign(PrAlignArgs a)
{
    const uint32_t item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= a.nitems) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t pair = a.items[item];
    const CharT* A = (const CharT*)a.achars + a.aoff[pair];
    const CharT* B = (const CharT*)a.bchars + a.boff[pair];
    const uint32_t la = (uint32_t)(a.aoff[pair + 1] - a.aoff[pair]), lb = (uint32_t)(a.boff[pair + 1] - a.boff[pair]);
    const bool swap = la > lb; // partial_ratio_alignment(s2, s1) with the fields exchanged
    const CharT* nd = swap ? B : A;
    const CharT* hy = swap ? A : B;
    const uint32_t l1 = swap ? lb : la, l2 = swap ? la : lb;
    uint8_t* base = a.scratch + a.scratch_off[item];
    const uint64_t W = (l1 + 63) / 64;
    uint64_t* xtab = (uint64_t*)base;
    uint64_t* xtab2 = xtab + W * l2;
    uint32_t* lcs = (uint32_t*)(xtab2 + (l1 == l2 ? W * l1 : 0));
    uint2* lists = (uint2*)(base + (pralign_scratch_bytes(l1, l2) - (2ull * (l2 - l1) + 2) * 2 * 8));

    PrResult r = pr_search<WM, CharT>(nd, l1, hy, l2, a.cutoff, xtab, lcs, lists);
    uint32_t ss = 0, se = l1, dst = r.ds, den = r.de;
    double score = r.score;
    if (score != 100.0 && l1 == l2) {
        const double c2 = a.cutoff < score ? score : a.cutoff; // std::max(score_cutoff, score)
        const PrResult r2 = pr_search<WM, CharT>(hy, l2, nd, l1, c2, xtab2, lcs, lists);
        if (r2.score > score) {
            score = r2.score;
            ss = r2.ds;
            se = r2.de;
            dst = 0;
            den = l2;
        }
    }
    if (lane == 0) {
        a.score[pair] = score;
        uint64_t* o = a.al + 4ull * pair;
        if (swap) {
            o[0] = dst;
            o[1] = den;
            o[2] = ss;
            o[3] = se;
        }
        else {
            o[0] = ss;
            o[1] = se;
            o[2] = dst;
            o[3] = den;
        }
    }
}

template <uint32_t WM>
static void launch_w(int char_width, const PrAlignArgs& a, uint32_t blocks, hipStream_t s)
{
    if (char_width == 1)
        hipLaunchKernelGGL((k_pralign<WM, uint8_t>), dim3(blocks), dim3(256), 0, s, a);
    else if (char_width == 2)
        hipLaunchKernelGGL((k_pralign<WM, uint16_t>), dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_pralign<WM, uint32_t>), dim3(blocks), dim3(256), 0, s, a);
}

hipError_t launch_pralign(int char_width, uint32_t words, const PrAlignArgs& a, hipStream_t s)
{
    if (a.nitems == 0) return hipSuccess;
    const uint32_t blocks = (a.nitems + 3) / 4;
    if (words <= 1)
        launch_w<1>(char_width, a, blocks, s);
    else if (words <= 4)
        launch_w<4>(char_width, a, blocks, s);
    else
        launch_w<16>(char_width, a, blocks, s);
    return hipGetLastError();
}

} // namespace rfa
