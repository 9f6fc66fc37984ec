// rfa_partial.hip -- fuzz::partial_ratio for cdist (fuzz_impl.hpp:57-279,
// CachedPartialRatio::similarity fuzz_impl.hpp:255-279).
//
// The reference scores the shorter string s1 (the needle, length l1) against
// windows of the longer one s2 (length l2 >= l1):
//   full windows   s2[p, p + l1) for p < l2 - l1, searched by bisection with a
//                  distance bound; their best Indel distance d counts only if
//                  d < ceil(2 l1 * NormSim_to_NormDist(cutoff / 100)), and then
//                  scores (1 - d / 2 l1) * 100 when that is >= cutoff;
//   prefix windows s2[0, i) for 0 < i < l1, skipped unless s2[i - 1] is in s1;
//   suffix windows s2[i, l2) for l2 - l1 <= i < l2, skipped unless s2[i] is in
//                  s1; both score CachedRatio(s1).similarity(window, cutoff);
//   and, when l1 == l2 and the score is below 100, the same with the roles
//   swapped.  The result is the largest score found (each score passes the
//   cutoff on its own: a window scoring above the running best also passes the
//   raised cutoff, and ratios of distinct fractions never round to the same
//   double), so every window can be scored independently.  The bisection's
//   pruning bound holds exactly (adjacent windows differ by at most 2 in Indel
//   distance, and distances of equal-length windows are even), so the minimum
//   over all full windows equals the one it finds whenever that passes.
//
// A prefix or suffix window's score is a function of x = d / M alone (d its
// Indel distance, M = l1 + L): the reference's test d <= ceil(M * c') rejects
// only windows that its test fl(x) <= c' rejects too (for M < 2^19, M * c'
// cannot fall a whole unit below d while x rounds to c'), and the other tests
// compare fl(x).  So the best such window is the one with the largest ratio
// 2 lcs / M; the kernels compare those fractions in integers and compute the
// doubles once per pair: that window's score and the full windows' score.
//
// Tile kernel (queries of at most 64 units, one per lane, PM[row][lane] in LDS,
// the choice's code units uniform across the wave):
//   query = needle (m <= n): for every window start p of the choice, the query's
//     PM against c[p + j], j < min(m, n - p); the window's LCS is popcount(~S).
//     Window 0's running counts give the prefix windows.
//   choice = needle (m >= n): the pattern of the window q[p, p + n) is the
//     query's PM shifted right by p and masked to n bits, the text is the
//     choice; the final state of window 0 gives every prefix window
//     (popcount of its low i bits).
// Long kernel (longer queries, one wave per pair): the same two cases with the
// windows across the lanes and the needle's words in registers.
#include <cmath>

#include "rfa_block_impl.hpp"
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

// best prefix / suffix window so far by its ratio num / den = 2 lcs / (l1 + L); num = 0: none
struct WBest {
    uint32_t num, den;
};

__device__ __forceinline__ void offer(WBest& b, uint32_t l1, uint32_t L, uint32_t lcs)
{
    const uint32_t num = 2 * lcs, den = l1 + L;
    if ((uint64_t)num * b.den > (uint64_t)b.num * den) {
        b.num = num;
        b.den = den;
    }
}

// CachedRatio(s1).similarity(window, cutoff) of the best window (fuzz_impl.hpp:43-47,
// 151-175): Indel distance den - num against the maximum den
__device__ __forceinline__ double window_score(const ScoreSpec& indel, const WBest& b, double cut)
{
    if (b.num == 0) return 0.0;
    IntPair p{&indel, b.den, 0, b.num / 2};
    return int_api_norm_similarity(p, cut / 100) * 100;
}

// the full windows' best distance d (fuzz_impl.hpp:81-148); 0 when it does not count
__device__ __forceinline__ double full_score(uint32_t l1, uint32_t d, double cut)
{
    if (d == 0xffffffffu) return 0.0;
    const uint64_t maximum = 2ull * l1;
    const double ncs = dmin(1.0, 1.0 - cut / 100 + 0.00001); // NormSim_to_NormDist
    const uint64_t cutoff_dist = (uint64_t)ceil((double)maximum * ncs);
    if ((uint64_t)d >= cutoff_dist) return 0.0;
    double score = 1.0 - (double)d / (double)maximum;
    score *= 100;
    return score >= cut ? score : 0.0;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return uniform(v);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off, 64));
    return uniform(v);
}

// one-word LCS state (LCSseq_impl.hpp:219-252); 64-bit words as two halves
template <typename W>
struct Lcs1;

template <>
struct Lcs1<uint32_t> {
    uint32_t S;
    __device__ __forceinline__ void reset() { S = ~0u; }
    __device__ __forceinline__ void step(uint32_t X) { lcs_step(X, S); }
    __device__ __forceinline__ uint32_t count() const { return (uint32_t)__popc(~S); }
    __device__ __forceinline__ uint32_t zero_bit(uint32_t k) const { return (~S >> k) & 1u; }
};

template <>
struct Lcs1<uint64_t> {
    uint32_t Sl, Sh;
    __device__ __forceinline__ void reset() { Sl = Sh = ~0u; }
    __device__ __forceinline__ void step(uint64_t X) { lcs_step(X, Sl, Sh); }
    __device__ __forceinline__ uint32_t count() const { return (uint32_t)(__popc(~Sl) + __popc(~Sh)); }
    __device__ __forceinline__ uint32_t zero_bit(uint32_t k) const
    {
        return k < 32 ? (~Sl >> k) & 1u : (~Sh >> (k - 32)) & 1u;
    }
};

// ---- tile kernel --------------------------------------------------------------

template <typename CharT, typename W>
__global__ __launch_bounds__(TILE_THREADS) void k_ptile(TileArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    W* pm = (W*)smem_u64;
    uint8_t* extra = (uint8_t*)(pm + (size_t)a.rows_cap * 64);
    const CharT* __restrict__ q = (const CharT*)a.qchars;
    const CharT* __restrict__ c = (const CharT*)a.cchars;
    constexpr uint32_t BITS = sizeof(W) * 8;

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tile = blockIdx.y + a.tile_base;
    const uint32_t qpos = tile * 64 + lane;
    const bool qact = qpos < a.nq;
    const uint32_t qidx = qact ? a.qorder[qpos] : 0;
    const uint64_t qo = qact ? a.qoff[qidx] : 0;
    const uint32_t m = qact ? (uint32_t)(a.qoff[qidx + 1] - qo) : 0;

    SymLookup<CharT> look = build_tile_pm<CharT, W>(q, qo, m, pm, extra, a);
    ScoreSpec indel = a.spec;
    indel.metric = RFA_INDEL;
    const double cut_all = a.spec.cutoff_d;
    const W* __restrict__ col = pm + lane;

    const uint32_t it0 = blockIdx.x * a.choices_per_block;
    const uint32_t it1 = min(a.ncsel, it0 + a.choices_per_block);
    constexpr uint32_t G = 4;
    for (uint32_t g0 = it0 + uniform(wave) * G; g0 < it1; g0 += TILE_WAVES * G) {
        const uint32_t gn = min(G, it1 - g0);
        uint64_t outv[G];
        for (uint32_t t = 0; t < gn; ++t) {
            const uint32_t cidx = g0 + t;
            const double cut = a.pcut && qact ? a.pcut[(uint64_t)qidx * a.out_stride + cidx] : cut_all;
            const uint64_t co0 = a.coff[cidx];
            const uint32_t n = uniform((uint32_t)(a.coff[cidx + 1] - co0));
            const uint64_t co = ((uint64_t)uniform((uint32_t)(co0 >> 32)) << 32) | uniform((uint32_t)co0);
            WBest b{0u, 1u};
            uint32_t bestd1 = 0xffffffffu, bestd2 = 0xffffffffu; // full windows of each case

            // ---- the query is the needle: windows of the choice ----
            const bool on1 = qact && m > 0 && m <= n && cut <= 100;
            const uint32_t mm = wave_max_u32(on1 ? m : 0u);
            if (mm) {
                // steps below the shortest participating query need no per-lane length mask
                const uint32_t mmin = 64 - wave_max_u32(on1 ? 64 - m : 0u);
                // one register holds the rows of c[k0, k0 + 64): windows k0 .. k0 + 64 - mm
                // never read past it
                const uint32_t wstep = 65 - mm;
                uint32_t unit_next = lane < n ? (uint32_t)c[co + lane] : 0u;
                for (uint32_t k0 = 0; k0 < n; k0 += wstep) {
                    const uint32_t row = k0 + lane < n ? look(unit_next) : 0u;
                    // the next batch's unit, loaded while this batch runs
                    const uint32_t k1 = k0 + wstep;
                    if (k1 + lane < n) unit_next = (uint32_t)c[co + k1 + lane];
                    const uint32_t pend = min(n, k0 + wstep);
                    uint32_t p = k0;
                    for (; p < pend; ++p) {
                        const uint32_t steps = min(mm, n - p);
                        const uint32_t free_steps = min(mmin, steps);
                        const uint32_t i0 = p - k0;
                        Lcs1<W> S;
                        S.reset();
                        // two rows in flight: Xa for step j, Xb for step j + 1
                        W Xa = col[readlane(row, i0) * 64];
                        W Xb = col[readlane(row, min(i0 + 1, 63u)) * 64];
                        const bool first_in = Xa != 0;
                        uint32_t j = 0;
                        for (; j + 2 <= free_steps; j += 2) {
                            const W A = Xa, B = Xb;
                            Xa = col[readlane(row, min(i0 + j + 2, 63u)) * 64];
                            Xb = col[readlane(row, min(i0 + j + 3, 63u)) * 64];
                            S.step(A);
                            S.step(B);
                        }
                        for (; j < steps; ++j) {
                            const W X = Xa;
                            Xa = Xb;
                            Xb = col[readlane(row, min(i0 + j + 2, 63u)) * 64];
                            S.step(j < m ? X : (W)0);
                        }
                        if (on1) {
                            const uint32_t lcs = S.count();
                            if (p + m < n)
                                bestd1 = min(bestd1, 2 * (m - lcs));
                            else if (first_in)
                                offer(b, m, n - p, lcs); // suffix window c[p, n)
                        }
                    }
                }
                // prefix windows c[0, i), 0 < i < m (each needs its last unit in the query):
                // window 0 again, step by step
                if (wave_max_u32(on1 && m > 1 ? 1u : 0u)) {
                    const uint32_t row = lane < n ? look((uint32_t)c[co + lane]) : 0u;
                    Lcs1<W> S;
                    S.reset();
                    const uint32_t steps = min(mm, n) - 1;
                    for (uint32_t j = 0; j < steps; ++j) {
                        W X = col[readlane(row, j) * 64];
                        X = j < m ? X : (W)0;
                        S.step(X);
                        if (on1 && j + 1 < m && X != 0) offer(b, m, j + 1, S.count());
                    }
                }
            }

            // ---- the choice is the needle (and the swapped pass of equal lengths) ----
            const bool on2 = qact && n > 0 && m >= n && cut <= 100;
            const uint32_t mM = wave_max_u32(on2 ? m : 0u);
            if (mM) {
                // n <= m <= 64: the whole choice in one register
                uint32_t rowC = 0;
                if (lane < n) rowC = look((uint32_t)c[co + lane]);
                const W maskn = n >= BITS ? ~(W)0 : (((W)1 << n) - 1);
                W U = 0; // bit k: q[k] occurs in the choice
                for (uint32_t j = 0; j < n; ++j) U |= col[readlane(rowC, j) * 64];
                Lcs1<W> S0; // final state of window 0
                S0.reset();
                for (uint32_t p = 0; p < mM; ++p) {
                    Lcs1<W> S;
                    S.reset();
                    W Xn = col[readlane(rowC, 0) * 64];
                    for (uint32_t j = 0; j < n; ++j) {
                        const W Xq = Xn;
                        Xn = col[readlane(rowC, min(j + 1, 63u)) * 64];
                        S.step((W)(Xq >> p) & maskn);
                    }
                    if (p == 0) S0 = S;
                    if (on2 && p < m) {
                        const uint32_t lcs = S.count();
                        if (p + n < m)
                            bestd2 = min(bestd2, 2 * (n - lcs));
                        else if ((U >> p) & 1)
                            offer(b, n, m - p, lcs); // suffix window q[p, m)
                    }
                }
                if (on2) {
                    // prefix windows q[0, i), 0 < i < n: LCS = zero bits of window 0's state below i
                    uint32_t cum = 0;
                    for (uint32_t i = 1; i < n; ++i) {
                        cum += S0.zero_bit(i - 1);
                        if ((U >> (i - 1)) & 1) offer(b, n, i, cum);
                    }
                }
            }

            double v;
            if (cut > 100)
                v = 0.0;
            else if (m == 0 || n == 0)
                v = (m == n) ? 100.0 : 0.0;
            else
                v = dmax(window_score(indel, b, cut), dmax(full_score(m, bestd1, cut), full_score(n, bestd2, cut)));
            union {
                double d;
                uint64_t u;
            } cv;
            cv.d = v;
            outv[t] = cv.u;
        }
        if (qact) {
            uint64_t* dst = (uint64_t*)a.out + (uint64_t)qidx * a.out_stride + g0;
            if (gn == G && ((uintptr_t)dst & 15) == 0) {
                uint4* d4 = (uint4*)dst;
#pragma unroll
                for (uint32_t t = 0; t < G; t += 2)
                    d4[t / 2] = make_uint4((uint32_t)outv[t], (uint32_t)(outv[t] >> 32), (uint32_t)outv[t + 1],
                                           (uint32_t)(outv[t + 1] >> 32));
            }
            else {
                for (uint32_t t = 0; t < gn; ++t) dst[t] = outv[t];
            }
        }
    }
}

// ---- long kernel: one wave per pair, windows across the lanes ------------------

// S += S & X with the carry chain across the needle's words; S | (S & ~X)
template <int WMAX>
__device__ __forceinline__ void lcs_words(uint64_t (&S)[WMAX], const uint64_t (&X)[WMAX], uint32_t nw)
{
    uint64_t carry = 0;
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
        if ((uint32_t)w < nw) {
            const uint64_t s = S[w], u = s & X[w];
            const uint64_t t = s + u;
            const uint64_t c1 = t < s;
            const uint64_t t2 = t + carry;
            const uint64_t c2 = t2 < t;
            S[w] = t2 | (s & ~X[w]);
            carry = c1 | c2;
        }
    }
}

template <int WMAX>
__device__ __forceinline__ uint32_t zeros(const uint64_t (&S)[WMAX], uint32_t nw)
{
    uint32_t z = 0;
#pragma unroll
    for (int w = 0; w < WMAX; ++w)
        if ((uint32_t)w < nw) z += (uint32_t)__popcll(~S[w]);
    return z;
}

// the query (m units) is the needle, m <= n: window p = lane + 64k of the choice
template <int WMAX, typename CharT, bool LDS>
__device__ void plong_query_needle(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ T, uint32_t n,
                                   WBest& b, uint32_t& bestd)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (m + 63) / 64;
    for (uint32_t p0 = 0; p0 < n; p0 += 64) {
        const uint32_t p = p0 + lane;
        const bool act = p < n;
        const uint32_t L = act ? min(m, n - p) : 0u;
        const uint32_t steps = min(m, n - p0);
        uint64_t S[WMAX], X[WMAX];
#pragma unroll
        for (int w = 0; w < WMAX; ++w) S[w] = ~0ull;
        bool first_in = false;
        for (uint32_t j = 0; j < steps; ++j) {
            const uint32_t r = j < L ? P.map.find((uint32_t)T[p + j]) : 0u;
            if (j == 0) first_in = r != 0;
            const uint64_t* row = P.pm + (size_t)r * P.Wp;
#pragma unroll
            for (int w = 0; w < WMAX; ++w) X[w] = (uint32_t)w < nw ? row[w] : 0ull;
            lcs_words<WMAX>(S, X, nw);
            // prefix window c[0, j + 1) (lane 0 holds window 0)
            if (p == 0 && j + 1 < m && r != 0) offer(b, m, j + 1, zeros<WMAX>(S, nw));
        }
        if (act) {
            const uint32_t lcs = zeros<WMAX>(S, nw);
            if (p + m < n)
                bestd = min(bestd, 2 * (m - lcs));
            else if (first_in)
                offer(b, m, n - p, lcs);
        }
    }
}

// the choice (n units) is the needle, n <= m: window q[p, p + n) with p = lane + 64k,
// its pattern the query's PM from bit p, masked to n bits; the text is the choice
template <int WMAX, typename CharT, bool LDS>
__device__ void plong_choice_needle(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ T, uint32_t n,
                                    WBest& b, uint32_t& bestd)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nw = (n + 63) / 64, qw = P.W;
    const uint64_t last = (n & 63) ? ((1ull << (n & 63)) - 1) : ~0ull;
    for (uint32_t p0 = 0; p0 < m; p0 += 64) {
        const uint32_t p = p0 + lane;
        const bool act = p < m;
        const uint32_t base = p >> 6, sh = p & 63;
        uint64_t S[WMAX], U[WMAX], X[WMAX];
#pragma unroll
        for (int w = 0; w < WMAX; ++w) {
            S[w] = ~0ull;
            U[w] = 0;
        }
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t r = P.map.find((uint32_t)T[j]);
            const uint64_t* row = P.pm + (size_t)r * P.Wp;
#pragma unroll
            for (int w = 0; w < WMAX; ++w) {
                uint64_t x = 0;
                if ((uint32_t)w < nw) {
                    const uint32_t k = base + w;
                    const uint64_t lo = (act && k < qw) ? row[k] : 0ull;
                    const uint64_t hi = (act && k + 1 < qw) ? row[k + 1] : 0ull;
                    x = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
                    if ((uint32_t)w == nw - 1) x &= last;
                    U[w] |= x;
                }
                X[w] = x;
            }
            lcs_words<WMAX>(S, X, nw);
        }
        if (act) {
            const uint32_t lcs = zeros<WMAX>(S, nw);
            if (p + n < m)
                bestd = min(bestd, 2 * (n - lcs));
            else if (U[0] & 1)
                offer(b, n, m - p, lcs); // suffix window q[p, m), q[p] in the choice
            if (p == 0) {
                // prefix windows q[0, i), 0 < i < n
                uint32_t cum = 0;
#pragma unroll
                for (int w = 0; w < WMAX; ++w) {
                    if ((uint32_t)w < nw) {
                        const uint64_t z = ~S[w], u = U[w];
                        for (uint32_t k = 0; k < 64; ++k) {
                            const uint32_t i = (uint32_t)w * 64 + k + 1;
                            if (i >= n) break;
                            cum += (uint32_t)((z >> k) & 1);
                            if ((u >> k) & 1) offer(b, n, i, cum);
                        }
                    }
                }
            }
        }
    }
}

template <typename CharT, int WMAX, bool LDS>
__global__ __launch_bounds__(64) void k_plong(PartialLongArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* arena = LDS ? (uint8_t*)smem_u64 : (uint8_t*)(a.pm_scratch + (size_t)blockIdx.x * a.pm_scratch_words);
    const CharT* __restrict__ Q = (const CharT*)a.qchars;
    const CharT* __restrict__ C = (const CharT*)a.cchars;
    const uint32_t lane = threadIdx.x & 63;
    ScoreSpec indel{};
    indel.metric = RFA_INDEL;
    indel.kind = RFA_NORMALIZED_SIMILARITY;
    // a contiguous share of the (query, choice) items: the pattern changes rarely
    const uint64_t total = (uint64_t)a.nql * a.nc;
    const uint64_t lo = total * blockIdx.x / gridDim.x, hi = total * (blockIdx.x + 1) / gridDim.x;
    uint32_t prev = 0xffffffffu;
    Pattern<CharT, LDS> P{};
    for (uint64_t it = lo; it < hi; ++it) {
        const uint32_t qk = (uint32_t)(it / a.nc), ci = (uint32_t)(it % a.nc);
        const uint32_t qi = a.qlist[qk];
        const uint64_t qo = a.qoff[qi];
        const uint32_t m = (uint32_t)(a.qoff[qi + 1] - qo);
        if (qi != prev) {
            P = build_pattern<CharT, LDS>(Q + qo, m, arena, pattern_hash_bits<CharT>(m));
            prev = qi;
        }
        const uint64_t co = a.coff[ci];
        const uint32_t n = (uint32_t)(a.coff[ci + 1] - co);
        if (min(m, n) > 64u * WMAX) continue; // a longer needle: k_pbig's pair
        const double cut = a.pcut ? a.pcut[(uint64_t)qi * a.out_stride + ci] : a.cutoff;
        double v;
        if (cut > 100) {
            v = 0.0;
        }
        else if (m == 0 || n == 0) {
            v = (m == n) ? 100.0 : 0.0;
        }
        else {
            WBest b{0u, 1u};
            uint32_t bestd1 = 0xffffffffu, bestd2 = 0xffffffffu;
            if (m <= n) plong_query_needle<WMAX>(P, m, C + co, n, b, bestd1);
            if (m >= n) plong_choice_needle<WMAX>(P, m, C + co, n, b, bestd2);
            // the wave's best window ratio and full-window distances
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t on = (uint32_t)__shfl_xor(b.num, off, 64), od = (uint32_t)__shfl_xor(b.den, off, 64);
                if ((uint64_t)on * b.den > (uint64_t)b.num * od) {
                    b.num = on;
                    b.den = od;
                }
            }
            bestd1 = wave_min_u32(bestd1);
            bestd2 = wave_min_u32(bestd2);
            v = dmax(window_score(indel, b, cut), dmax(full_score(m, bestd1, cut), full_score(n, bestd2, cut)));
        }
        if (lane == 0) a.out[(uint64_t)qi * a.out_stride + ci] = v;
    }
}

// ---- launches -------------------------------------------------------------------

template <typename CharT, typename W>
static void launch_ptile_w(const TileArgs& a, uint32_t ntiles, uint32_t nbx, size_t lds, hipStream_t s)
{
    hipLaunchKernelGGL((k_ptile<CharT, W>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a);
}

hipError_t launch_ptile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                        hipStream_t s)
{
    const size_t lds = tile_lds_bytes(char_width, word_bits, a.rows_cap, a.hash_bits);
    if (word_bits == 32) {
        if (char_width == 1)
            launch_ptile_w<uint8_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_ptile_w<uint16_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_ptile_w<uint32_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
    }
    else {
        if (char_width == 1)
            launch_ptile_w<uint8_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_ptile_w<uint16_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_ptile_w<uint32_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
    }
    return hipGetLastError();
}

template <typename CharT, int WMAX>
static void launch_plong_w(const PartialLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (a.lds_bytes)
        hipLaunchKernelGGL((k_plong<CharT, WMAX, true>), dim3(nblocks), dim3(64), a.lds_bytes, s, a);
    else
        hipLaunchKernelGGL((k_plong<CharT, WMAX, false>), dim3(nblocks), dim3(64), 0, s, a);
}

template <typename CharT>
static void launch_plong_c(uint32_t words, const PartialLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (words <= 1)
        launch_plong_w<CharT, 1>(a, nblocks, s);
    else if (words <= 2)
        launch_plong_w<CharT, 2>(a, nblocks, s);
    else if (words <= 4)
        launch_plong_w<CharT, 4>(a, nblocks, s);
    else if (words <= 8)
        launch_plong_w<CharT, 8>(a, nblocks, s);
    else
        launch_plong_w<CharT, 16>(a, nblocks, s);
}

hipError_t launch_plong(int char_width, uint32_t needle_words, const PartialLongArgs& a, uint32_t nblocks,
                        hipStream_t s)
{
    if (needle_words > PARTIAL_MAX_WORDS) return hipErrorInvalidValue;
    if (char_width == 1)
        launch_plong_c<uint8_t>(needle_words, a, nblocks, s);
    else if (char_width == 2)
        launch_plong_c<uint16_t>(needle_words, a, nblocks, s);
    else
        launch_plong_c<uint32_t>(needle_words, a, nblocks, s);
    return hipGetLastError();
}

} // namespace rfa
