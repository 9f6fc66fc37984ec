// rfa_tset.hip -- fuzz::token_set_ratio and the set part of fuzz::token_ratio
// for cdist (fuzz_impl.hpp:370-426, 518-565; set_decomposition common_impl.hpp:11-35).
//
// Both strings are taken as their deduplicated, sorted token lists U (joined
// with single spaces) plus a token table (start in U, length, hash).  Per pair:
//   sect = tokens in both, ab = the query's other tokens, ba = the choice's;
//   dist = indel_distance(join(ab), join(ba)).
// join(ab) is U_q with the sect tokens (and the spaces they leave doubled)
// removed; a unit that can match nothing does not change an LCS, so
// LCS(join(ab), join(ba)) is the bit-parallel LCS of U_q's pattern with the
// removed positions masked out against U_c with the removed units stepped as
// "no match".  No per-pair string is built.
//
// partial_token_set_ratio and partial_token_ratio (fuzz_impl.hpp:460-478, 740-770)
// return 100 when the sets share a token; otherwise their differences are the
// whole deduplicated sets, so they are partial_ratio of the deduplicated joins
// (and, for partial_token_ratio, of the sorted joins: the larger, the two being
// equal when no token repeats).  The host computes those with the partial_ratio
// path; modes 2 and 3 here only decide "shares a token".
//
// Query tiles (the scorer's tile plan over the queries' U, at most 64 units,
// one query per lane, PM[row][lane] in LDS) against choices uniform across the
// wave.  Each lane keeps its query's token table in LDS columns and a 64-bit
// Bloom mask of its token hashes; a choice token is compared only when the
// Bloom bit is set, and tokens whose hashes match are compared unit by unit.
#include <cmath>

#include "rfa_block_impl.hpp"
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

constexpr uint32_t TSET_MAX_QTOK = 32; // a join of <= 64 units has <= 32 tokens

// fuzz_detail::norm_distance (fuzz_impl.hpp:55-61)
__device__ __forceinline__ double fz_norm_distance(uint64_t dist, uint64_t lensum, double cut)
{
    const double score = (lensum > 0) ? (100.0 - 100.0 * (double)dist / (double)lensum) : 100.0;
    return (score >= cut) ? score : 0.0;
}

template <typename W>
struct TsLcs;

template <>
struct TsLcs<uint32_t> {
    uint32_t S;
    __device__ __forceinline__ void reset() { S = ~0u; }
    __device__ __forceinline__ void step(uint32_t X) { lcs_step(X, S); }
    __device__ __forceinline__ uint32_t count() const { return (uint32_t)__popc(~S); }
};

template <>
struct TsLcs<uint64_t> {
    uint32_t Sl, Sh;
    __device__ __forceinline__ void reset() { Sl = Sh = ~0u; }
    __device__ __forceinline__ void step(uint64_t X) { lcs_step(X, Sl, Sh); }
    __device__ __forceinline__ uint32_t count() const { return (uint32_t)(__popc(~Sl) + __popc(~Sh)); }
};

template <typename CharT, typename W, int MODE>
__global__ __launch_bounds__(TILE_THREADS) void k_tset(TsetArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    const TileArgs& ta = a.t;
    W* pm = (W*)smem_u64;
    uint8_t* extra = (uint8_t*)(pm + (size_t)ta.rows_cap * 64);
    // the query token tables after the tile's PM rows and symbol map
    uint64_t* qhash = (uint64_t*)(smem_u64 + a.tok_lds_u64);
    uint8_t* qst = (uint8_t*)(qhash + TSET_MAX_QTOK * 64);
    uint8_t* qln = qst + TSET_MAX_QTOK * 64;
    const CharT* __restrict__ q = (const CharT*)ta.qchars;
    const CharT* __restrict__ c = (const CharT*)ta.cchars;

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tile = blockIdx.y + ta.tile_base;
    const uint32_t qpos = tile * 64 + lane;
    const bool qact = qpos < ta.nq;
    const uint32_t qidx = qact ? ta.qorder[qpos] : 0;
    const uint64_t qo = qact ? ta.qoff[qidx] : 0;
    const uint32_t m = qact ? (uint32_t)(ta.qoff[qidx + 1] - qo) : 0;

    SymLookup<CharT> look = build_tile_pm<CharT, W>(q, qo, m, pm, extra, ta);
    const W* __restrict__ col = pm + lane;
    const double cut_all = ta.spec.cutoff_d;

    // this lane's query tokens: hash, start and length in LDS columns (each wave holds
    // the same tile, so every wave writes the same values into its own lane's column)
    uint32_t Tq = 0;
    uint64_t bloom = 0;
    if (qact) {
        const uint64_t t0 = a.q_set_off[qidx];
        Tq = (uint32_t)(a.q_set_off[qidx + 1] - t0);
        for (uint32_t k = 0; k < Tq; ++k) {
            const uint64_t h = a.q_set_hash[t0 + k];
            if (wave == 0) {
                qhash[k * 64 + lane] = h;
                qst[k * 64 + lane] = (uint8_t)a.q_set_start[t0 + k];
                qln[k * 64 + lane] = (uint8_t)a.q_set_len[t0 + k];
            }
            bloom |= 1ull << (h & 63);
        }
    }
    __syncthreads();

    // is choice token (h, len, units at p) one of this lane's query tokens?  -1: no, else its index
    auto find_q = [&](uint64_t h, uint32_t len, const CharT* __restrict__ p) -> int {
        if (!qact || !((bloom >> (h & 63)) & 1)) return -1;
        for (uint32_t k = 0; k < Tq; ++k) {
            if (qhash[k * 64 + lane] != h || qln[k * 64 + lane] != len) continue;
            const CharT* qs = q + qo + qst[k * 64 + lane];
            uint32_t i = 0;
            while (i < len && qs[i] == p[i]) ++i;
            if (i == len) return (int)k;
        }
        return -1;
    };

    const uint32_t it0 = blockIdx.x * ta.choices_per_block;
    const uint32_t it1 = min(ta.ncsel, it0 + ta.choices_per_block);
    for (uint32_t cidx = it0 + uniform(wave); cidx < it1; cidx += TILE_WAVES) {
        const uint64_t co0 = ta.coff[cidx];
        const uint64_t co = ((uint64_t)uniform((uint32_t)(co0 >> 32)) << 32) | uniform((uint32_t)co0);
        const uint64_t tc0 = a.c_set_off[cidx];
        const uint32_t Tc = uniform((uint32_t)(a.c_set_off[cidx + 1] - tc0));
        const double cut = ta.pcut && qact ? ta.pcut[(uint64_t)qidx * ta.out_stride + cidx] : cut_all;
        double v = 0.0;
        bool perfect = false, common = false;
        if (ballot(cut <= 100)) {
            // ---- pass 1: the intersection ----
            uint32_t inter_q = 0, n_inter = 0, n_kept_c = 0;
            uint64_t ba_units = 0, kept_c_lo = 0;
            for (uint32_t t = 0; t < Tc; ++t) {
                const uint64_t h = a.c_set_hash[tc0 + t];
                const uint32_t st = a.c_set_start[tc0 + t], ln = a.c_set_len[tc0 + t];
                const int k = find_q(h, ln, c + co + st);
                if (k >= 0) {
                    inter_q |= 1u << k;
                    ++n_inter;
                }
                else {
                    ba_units += ln;
                    ++n_kept_c;
                    if (t < 64) kept_c_lo |= 1ull << t;
                }
            }
            // ---- the query side: ab, sect and the pattern mask of join(ab) ----
            uint64_t ab_units = 0, sect_units = 0;
            uint32_t n_kept_q = 0;
            W mask = 0;
            for (uint32_t k = 0; k < Tq; ++k) {
                const uint32_t st = qst[k * 64 + lane], ln = qln[k * 64 + lane];
                if ((inter_q >> k) & 1) {
                    sect_units += ln;
                }
                else {
                    ab_units += ln;
                    const W bits = (ln >= sizeof(W) * 8) ? ~(W)0 : (((W)1 << ln) - 1);
                    mask |= bits << st;
                    if (n_kept_q) mask |= (W)1 << (st - 1); // the space before it
                    ++n_kept_q;
                }
            }
            const uint64_t ab_len = ab_units + (n_kept_q ? n_kept_q - 1 : 0);
            const uint64_t ba_len = ba_units + (n_kept_c ? n_kept_c - 1 : 0);
            const uint64_t sect_len = sect_units + (n_inter ? n_inter - 1 : 0);
            perfect = n_inter > 0 && (n_kept_q == 0 || n_kept_c == 0);
            common = n_inter > 0;
            // ---- pass 2: LCS(join(ab), join(ba)) over U_c (token_set_ratio / token_ratio) ----
            const bool need = MODE < 2 && qact && !perfect;
            uint32_t lcs = 0;
            if (ballot(need)) {
                TsLcs<W> S;
                S.reset();
                bool seen = false;
                for (uint32_t t = 0; t < Tc; ++t) {
                    const uint32_t st = a.c_set_start[tc0 + t], ln = a.c_set_len[tc0 + t];
                    bool kept;
                    if (t < 64)
                        kept = (kept_c_lo >> t) & 1;
                    else
                        kept = find_q(a.c_set_hash[tc0 + t], ln, c + co + st) < 0;
                    if (t > 0) {
                        const W X = col[look((uint32_t)c[co + st - 1]) * 64];
                        S.step(kept && seen ? (X & mask) : (W)0);
                    }
                    for (uint32_t i = 0; i < ln; ++i) {
                        const W X = col[look((uint32_t)c[co + st + i]) * 64];
                        S.step(kept ? (X & mask) : (W)0);
                    }
                    seen = seen || kept;
                }
                lcs = S.count();
            }
            if (need) {
                const uint64_t dist = ab_len + ba_len - 2 * (uint64_t)lcs;
                const uint64_t sect_ab_len = sect_len + (sect_len != 0) + ab_len;
                const uint64_t sect_ba_len = sect_len + (sect_len != 0) + ba_len;
                const uint64_t lensum = sect_ab_len + sect_ba_len;
                const uint64_t cutoff_distance = (uint64_t)ceil((double)lensum * (1.0 - cut / 100));
                if (dist <= cutoff_distance) v = fz_norm_distance(dist, lensum, cut);
                if (sect_len) {
                    v = dmax(v, fz_norm_distance((sect_len != 0) + ab_len, sect_len + sect_ab_len, cut));
                    v = dmax(v, fz_norm_distance((sect_len != 0) + ba_len, sect_len + sect_ba_len, cut));
                }
            }
        }
        if (qact) {
            double* dst = (double*)ta.out + (uint64_t)qidx * ta.out_stride + cidx;
            double r;
            if (cut > 100)
                r = 0.0;
            else if (MODE == 0)
                r = (Tq == 0 || Tc == 0) ? 0.0 : (perfect ? 100.0 : v);
            else if (MODE == 1)
                r = perfect ? 100.0 : dmax(*dst, v); // token_ratio: out holds the sorted-join ratio
            else if (MODE == 2)
                r = (Tq == 0 || Tc == 0) ? 0.0 : (common ? 100.0 : *dst); // out: partial_ratio of the sets
            else
                r = common ? 100.0 : dmax(*dst, a.aux[(uint64_t)qidx * ta.out_stride + cidx]);
            *dst = r;
        }
    }
}

// ---- long queries: one wave per (query, choice) -------------------------------
//
// The query's join U_q (65..4096 units) is a pattern of W <= 64 words, word w in
// lane w, its PM rows built once per query into the workgroup's arena (LDS or
// scratch, build_pattern).  Per choice:
//   1. the query's tokens, 64 at a time one per lane, looked up among the choice's
//      (hash, length, then units): sect and ab sums, and the mask of join(ab) in U_q
//      (a kept token's units and, after the first kept token, the space before it),
//      set bit by bit into an LDS word per lane;
//   2. the choice's tokens the same way against the query's: ba sums, kept counts;
//   3. (token_set_ratio / token_ratio, unless settled) LCS(join(ab), join(ba)): U_c
//      streamed unit by unit, removed units stepping with no match, the kept units'
//      PM words masked; the words' additions carry across lanes by lookahead on two
//      ballots;
//   4. the ratios as the tile kernel forms them.

template <typename CharT>
__device__ __forceinline__ bool ts_tok_eq(const CharT* __restrict__ x, const CharT* __restrict__ y, uint32_t len)
{
    for (uint32_t i = 0; i < len; ++i)
        if (x[i] != y[i]) return false;
    return true;
}

template <typename CharT, bool LDS, int MODE>
__global__ __launch_bounds__(64) void k_tset_long(TsetLongArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    __shared__ uint64_t abmask[64]; // join(ab)'s units in U_q, word w at [w]
    const PartialLongArgs& pa = a.p;
    uint8_t* arena = LDS ? (uint8_t*)smem_u64 : (uint8_t*)(pa.pm_scratch + (size_t)blockIdx.x * pa.pm_scratch_words);
    const CharT* __restrict__ Q = (const CharT*)pa.qchars;
    const CharT* __restrict__ C = (const CharT*)pa.cchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t total = (uint64_t)pa.nql * pa.nc;
    const uint64_t lo = total * blockIdx.x / gridDim.x, hi = total * (blockIdx.x + 1) / gridDim.x;
    uint32_t prev = 0xffffffffu;
    Pattern<CharT, LDS> P{};
    for (uint64_t it = lo; it < hi; ++it) {
        const uint32_t qk = (uint32_t)(it / pa.nc), ci = (uint32_t)(it % pa.nc);
        const uint32_t qi = pa.qlist[qk];
        const uint64_t qo = pa.qoff[qi];
        const uint32_t m = uniform((uint32_t)(pa.qoff[qi + 1] - qo));
        if (qi != prev) {
            P = build_pattern<CharT, LDS>(Q + qo, m, arena, pattern_hash_bits<CharT>(m));
            prev = qi;
        }
        const uint64_t co = pa.coff[ci];
        const double cut = pa.pcut ? pa.pcut[(uint64_t)qi * pa.out_stride + ci] : pa.cutoff;
        const uint64_t tq0 = a.q_set_off[qi], tc0 = a.c_set_off[ci];
        const uint32_t Tq = uniform((uint32_t)(a.q_set_off[qi + 1] - tq0));
        const uint32_t Tc = uniform((uint32_t)(a.c_set_off[ci + 1] - tc0));
        double v = 0.0;
        bool perfect = false, common = false;
        if (cut <= 100) {
            // ---- 1. the query's tokens ----
            abmask[lane] = 0;
            __syncthreads();
            uint64_t ab_units = 0, sect_units = 0;
            uint32_t n_kept_q = 0, n_inter = 0, first_kept = 0xffffffffu;
            for (uint32_t k0 = 0; k0 < Tq; k0 += 64) {
                const uint32_t k = k0 + lane;
                const bool in = k < Tq;
                const uint32_t st = in ? a.q_set_start[tq0 + k] : 0u, ln = in ? a.q_set_len[tq0 + k] : 0u;
                bool found = false;
                if (in) {
                    const uint64_t h = a.q_set_hash[tq0 + k];
                    for (uint32_t t = 0; t < Tc && !found; ++t)
                        found = a.c_set_hash[tc0 + t] == h && a.c_set_len[tc0 + t] == ln &&
                                ts_tok_eq(Q + qo + st, C + co + a.c_set_start[tc0 + t], ln);
                }
                const bool kept = in && !found;
                const uint64_t km = ballot(kept);
                if (first_kept == 0xffffffffu && km) first_kept = k0 + (uint32_t)__builtin_ctzll(km);
                n_kept_q += (uint32_t)__popcll(km);
                n_inter += (uint32_t)__popcll(ballot(found));
                if (kept) {
                    ab_units += ln;
                    // the token's units, and the space before it after the first kept token
                    const uint32_t b0 = k != first_kept ? st - 1 : st, b1 = st + ln;
                    for (uint32_t b = b0; b < b1;) {
                        const uint32_t w = b / 64, lo_b = b % 64, nb = min(64u - lo_b, b1 - b);
                        const uint64_t bits = (nb == 64 ? ~0ull : ((1ull << nb) - 1)) << lo_b;
                        atomicOr((unsigned long long*)&abmask[w], (unsigned long long)bits);
                        b += nb;
                    }
                }
                else if (found) {
                    sect_units += ln;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                ab_units += (uint64_t)__shfl_xor((long long)ab_units, off, 64);
                sect_units += (uint64_t)__shfl_xor((long long)sect_units, off, 64);
            }
            // ---- 2. the choice's tokens ----
            uint64_t ba_units = 0;
            uint32_t n_kept_c = 0;
            for (uint32_t t0 = 0; t0 < Tc; t0 += 64) {
                const uint32_t t = t0 + lane;
                const bool in = t < Tc;
                bool found = false;
                uint32_t ln = 0;
                if (in) {
                    const uint64_t h = a.c_set_hash[tc0 + t];
                    const uint32_t st = a.c_set_start[tc0 + t];
                    ln = a.c_set_len[tc0 + t];
                    for (uint32_t k = 0; k < Tq && !found; ++k)
                        found = a.q_set_hash[tq0 + k] == h && a.q_set_len[tq0 + k] == ln &&
                                ts_tok_eq(C + co + st, Q + qo + a.q_set_start[tq0 + k], ln);
                }
                const bool kept = in && !found;
                n_kept_c += (uint32_t)__popcll(ballot(kept));
                if (kept) ba_units += ln;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) ba_units += (uint64_t)__shfl_xor((long long)ba_units, off, 64);
            const uint64_t ab_len = ab_units + (n_kept_q ? n_kept_q - 1 : 0);
            const uint64_t ba_len = ba_units + (n_kept_c ? n_kept_c - 1 : 0);
            const uint64_t sect_len = sect_units + (n_inter ? n_inter - 1 : 0);
            perfect = n_inter > 0 && (n_kept_q == 0 || n_kept_c == 0);
            common = n_inter > 0;
            // ---- 3. LCS(join(ab), join(ba)) ----
            if (MODE < 2 && !perfect) {
                __syncthreads();
                const uint32_t W = P.W;
                const bool act = lane < W;
                const uint64_t mw = act ? abmask[lane] : 0ull;
                uint64_t S = ~0ull;
                bool seen = false;
                auto step = [&](uint32_t unit, bool use) {
                    uint64_t X = 0;
                    if (use) {
                        const uint32_t r = P.map.find(unit);
                        X = act ? (P.pm[(size_t)r * P.Wp + lane] & mw) : 0ull;
                    }
                    const uint64_t u = S & X;
                    uint64_t x = S + u;
                    const bool g = act && (x < S);
                    const bool pr = act && !g && (x == ~0ull);
                    const uint64_t G = ballot(g), Pp = ballot(pr);
                    const uint64_t sum = (G | Pp) + G;
                    x += ((sum ^ Pp) >> lane) & 1u;
                    if (act) S = x | (S - u);
                };
                for (uint32_t t0 = 0; t0 < Tc; t0 += 64) {
                    const uint32_t t = t0 + lane;
                    const bool in = t < Tc;
                    bool found = false;
                    uint32_t st = 0, ln = 0;
                    if (in) {
                        const uint64_t h = a.c_set_hash[tc0 + t];
                        st = a.c_set_start[tc0 + t];
                        ln = a.c_set_len[tc0 + t];
                        for (uint32_t k = 0; k < Tq && !found; ++k)
                            found = a.q_set_hash[tq0 + k] == h && a.q_set_len[tq0 + k] == ln &&
                                    ts_tok_eq(C + co + st, Q + qo + a.q_set_start[tq0 + k], ln);
                    }
                    const uint64_t keptm = ballot(in && !found);
                    const uint32_t cnt = min(64u, Tc - t0);
                    for (uint32_t l = 0; l < cnt; ++l) {
                        const bool kept = (keptm >> l) & 1u;
                        const uint32_t tst = readlane(st, l), tln = readlane(ln, l);
                        if (t0 + l > 0) step((uint32_t)C[co + tst - 1], kept && seen);
                        for (uint32_t i = 0; i < tln; ++i) step((uint32_t)C[co + tst + i], kept);
                        seen = seen || kept;
                    }
                }
                const uint64_t valid = !act ? 0ull : (m - 64ull * lane >= 64 ? ~0ull : ((1ull << (m - 64ull * lane)) - 1));
                uint32_t lcs = (uint32_t)__popcll(~S & valid);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) lcs += (uint32_t)__shfl_xor((int)lcs, off, 64);
                const uint64_t dist = ab_len + ba_len - 2 * (uint64_t)lcs;
                const uint64_t sect_ab_len = sect_len + (sect_len != 0) + ab_len;
                const uint64_t sect_ba_len = sect_len + (sect_len != 0) + ba_len;
                const uint64_t lensum = sect_ab_len + sect_ba_len;
                const uint64_t cutoff_distance = (uint64_t)ceil((double)lensum * (1.0 - cut / 100));
                if (dist <= cutoff_distance) v = fz_norm_distance(dist, lensum, cut);
                if (sect_len) {
                    v = dmax(v, fz_norm_distance((sect_len != 0) + ab_len, sect_len + sect_ab_len, cut));
                    v = dmax(v, fz_norm_distance((sect_len != 0) + ba_len, sect_len + sect_ba_len, cut));
                }
            }
            __syncthreads(); // abmask is cleared for the next item
        }
        if (lane == 0) {
            double* dst = pa.out + (uint64_t)qi * pa.out_stride + ci;
            double r;
            if (cut > 100)
                r = 0.0;
            else if (MODE == 0)
                r = (Tq == 0 || Tc == 0) ? 0.0 : (perfect ? 100.0 : v);
            else if (MODE == 1)
                r = perfect ? 100.0 : dmax(*dst, v);
            else if (MODE == 2)
                r = (Tq == 0 || Tc == 0) ? 0.0 : (common ? 100.0 : *dst);
            else
                r = common ? 100.0 : dmax(*dst, a.aux[(uint64_t)qi * pa.out_stride + ci]);
            *dst = r;
        }
    }
}

template <typename CharT, bool LDS>
static void launch_tset_long_l(const TsetLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    const size_t lds = LDS ? a.p.lds_bytes : 0;
    switch (a.mode) {
    case 0: hipLaunchKernelGGL((k_tset_long<CharT, LDS, 0>), dim3(nblocks), dim3(64), lds, s, a); break;
    case 1: hipLaunchKernelGGL((k_tset_long<CharT, LDS, 1>), dim3(nblocks), dim3(64), lds, s, a); break;
    case 2: hipLaunchKernelGGL((k_tset_long<CharT, LDS, 2>), dim3(nblocks), dim3(64), lds, s, a); break;
    default: hipLaunchKernelGGL((k_tset_long<CharT, LDS, 3>), dim3(nblocks), dim3(64), lds, s, a); break;
    }
}

hipError_t launch_tset_long(int char_width, const TsetLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    const bool lds = a.p.lds_bytes != 0;
    if (char_width == 1)
        lds ? launch_tset_long_l<uint8_t, true>(a, nblocks, s) : launch_tset_long_l<uint8_t, false>(a, nblocks, s);
    else if (char_width == 2)
        lds ? launch_tset_long_l<uint16_t, true>(a, nblocks, s) : launch_tset_long_l<uint16_t, false>(a, nblocks, s);
    else
        lds ? launch_tset_long_l<uint32_t, true>(a, nblocks, s) : launch_tset_long_l<uint32_t, false>(a, nblocks, s);
    return hipGetLastError();
}

size_t tset_lds_bytes(int char_width, int word_bits, uint32_t rows_cap, uint32_t hash_bits, uint32_t* tok_lds_u64)
{
    const size_t tile = (tile_lds_bytes(char_width, word_bits, rows_cap, hash_bits) + 7) & ~(size_t)7;
    *tok_lds_u64 = (uint32_t)(tile / 8);
    return tile + TSET_MAX_QTOK * 64 * (8 + 2);
}

template <typename CharT, typename W>
static void launch_tset_w(const TsetArgs& a, uint32_t ntiles, uint32_t nbx, size_t lds, hipStream_t s)
{
    switch (a.mode) {
    case 0: hipLaunchKernelGGL((k_tset<CharT, W, 0>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a); break;
    case 1: hipLaunchKernelGGL((k_tset<CharT, W, 1>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a); break;
    case 2: hipLaunchKernelGGL((k_tset<CharT, W, 2>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a); break;
    default: hipLaunchKernelGGL((k_tset<CharT, W, 3>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a); break;
    }
}

// dst[q * stride + cols[j]] = src[q * ncols + j] for every q < nq, j < ncols
__global__ __launch_bounds__(256) void k_scatter_cols(double* dst, uint64_t stride, const double* src,
                                                      const uint32_t* cols, uint64_t ncols, uint64_t nq)
{
    const uint64_t total = nq * ncols;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t q = i / ncols, j = i % ncols;
        dst[q * stride + cols[j]] = src[i];
    }
}

hipError_t launch_scatter_cols(double* dst, uint64_t stride, const double* src, const uint32_t* cols, uint64_t ncols,
                               uint64_t nq, hipStream_t s)
{
    const uint64_t total = nq * ncols;
    if (!total) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_scatter_cols, dim3(blocks), dim3(256), 0, s, dst, stride, src, cols, ncols, nq);
    return hipGetLastError();
}

// dst[rows[r] * stride + c] = src[r * stride + c] for every r < nrows, c < stride
__global__ __launch_bounds__(256) void k_scatter_rows(double* dst, uint64_t stride, const double* src,
                                                      const uint32_t* rows, uint64_t nrows)
{
    const uint64_t total = nrows * stride;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / stride, c = i % stride;
        dst[rows[r] * stride + c] = src[i];
    }
}

hipError_t launch_scatter_rows(double* dst, uint64_t stride, const double* src, const uint32_t* rows, uint64_t nrows,
                               hipStream_t s)
{
    const uint64_t total = nrows * stride;
    if (!total) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_scatter_rows, dim3(blocks), dim3(256), 0, s, dst, stride, src, rows, nrows);
    return hipGetLastError();
}

hipError_t launch_tset(int char_width, int word_bits, TsetArgs a, uint32_t ntiles, uint32_t nblocks_x,
                       hipStream_t s)
{
    const size_t lds = tset_lds_bytes(char_width, word_bits, a.t.rows_cap, a.t.hash_bits, &a.tok_lds_u64);
    if (word_bits == 32) {
        if (char_width == 1)
            launch_tset_w<uint8_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_tset_w<uint16_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_tset_w<uint32_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
    }
    else {
        if (char_width == 1)
            launch_tset_w<uint8_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_tset_w<uint16_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_tset_w<uint32_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
    }
    return hipGetLastError();
}

} // namespace rfa
