// rfa_block_impl.hpp -- block kernel: one wave per (pattern, text) pair, the
// pattern's 64-bit words spread across the 64 lanes (word w in lane w % 64,
// register slot w / 64).
//
// The reference walks the words of one text position sequentially and passes
// the horizontal carries from word to word (Levenshtein_impl.hpp:689-743,
// LCSseq_impl.hpp:286-314, OSA_impl.hpp:199-243).  Here all words of a slot
// advance together and the carries are resolved with a carry-lookahead over
// the wave:
//
//   * Levenshtein / OSA: a word's outgoing HN carry depends only on its
//     incoming HN carry (one bit) and is monotone in it, so each lane
//     evaluates D0 for both values, publishes generate/propagate bits with
//     two ballots, and the scalar unit resolves all incoming carries at once
//     with one 64-bit add: carry_in = ((G|P) + G + c0) ^ P.  The HP carry is
//     then a one-lane shift of a third ballot.
//   * LCS: the same lookahead on the word additions S + (S & M).
//   * OSA additionally needs the previous word's D0/PM bit 63 for the
//     transposition term, again a shifted ballot.
//   * Jaro: the lowest matching, unflagged pattern position in the window is
//     the lowest set bit of the lowest lane with a nonzero masked word (ballot
//     + find-first-set), then the transposition walk pairs flags in order.
//
// Text code units are uniform across the wave, so every PM row read is one
// contiguous 64-lane access.  PM rows live in LDS when they fit, otherwise in
// a per-workgroup global arena.  Work items are pulled from a device queue by
// persistent single-wave workgroups (items are sorted longest first by the
// host, so the queue is a longest-processing-time-first schedule).
#pragma once
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"

namespace rfa {

// uint64 carry-out of a + b + c (c in {0,1}); all operands are wave-uniform.
__device__ __forceinline__ uint64_t add3_carry(uint64_t a, uint64_t b, uint64_t c, uint64_t* sum)
{
    uint64_t t = a + b;
    uint64_t c1 = t < a;
    uint64_t s = t + c;
    uint64_t c2 = s < t;
    *sum = s;
    return c1 | c2;
}

__device__ __forceinline__ uint32_t lane_bit(uint64_t mask, uint32_t lane) { return (uint32_t)(mask >> lane) & 1u; }

// Symbol -> PM row map of one pattern (per wave, in the arena).
template <typename CharT, bool LDS>
struct ArenaMap;

template <bool LDS>
struct ArenaMap<uint8_t, LDS> {
    uint32_t* rowmap; // [256]
    __device__ __forceinline__ uint32_t find(uint32_t c) const { return rowmap[c]; }
};

template <typename CharT, bool LDS>
struct ArenaMap {
    uint64_t* keys;
    uint32_t* rows;
    uint32_t bits;
    __device__ __forceinline__ uint32_t find(uint32_t c) const
    {
        uint32_t mask = (1u << bits) - 1;
        uint32_t h = sym_hash(c, bits);
        for (;;) {
            uint64_t k = keys[h];
            if (k == (uint64_t)c) return rows[h];
            if (k == EMPTY_KEY) return 0;
            h = (h + 1) & mask;
        }
    }
};

// Barrier for arena phases.  A global arena is written with plain stores and
// L2 atomics and then read back, so it also needs an agent-scope fence (the
// acquire side invalidates the CU's vector L1).
template <bool LDS>
__device__ __forceinline__ void arena_sync()
{
    __syncthreads();
    if constexpr (!LDS) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
}

// Wave-wide exclusive scan of v (one value per lane).
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* total)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    *total = __shfl(x, 63, 64);
    return x - v;
}

template <typename CharT, bool LDS>
struct Pattern {
    ArenaMap<CharT, LDS> map;
    uint64_t* pm; // [R][Wp], Wp = W (reads of lanes past the last word are masked)
    uint32_t R, W, Wp, slots;
};

// Builds the pattern's PM rows into the arena.  The arena layout is
// [map][pm]; map is 256 uint32 rows for 8-bit units, else 2^bits keys + rows.
template <typename CharT, bool LDS>
__device__ Pattern<CharT, LDS> build_pattern(const CharT* __restrict__ p, uint32_t m, uint8_t* arena,
                                              uint32_t hash_bits)
{
    const uint32_t lane = threadIdx.x & 63;
    Pattern<CharT, LDS> P;
    P.W = (m + 63) / 64;
    P.slots = (P.W + 63) / 64;
    P.Wp = P.W;
    uint64_t* pm;
    if constexpr (sizeof(CharT) == 1) {
        uint32_t* rowmap = (uint32_t*)arena;
        for (uint32_t i = lane; i < 256; i += 64) rowmap[i] = 0;
        arena_sync<LDS>();
        for (uint32_t i = lane; i < m; i += 64) rowmap[p[i]] = 1;
        arena_sync<LDS>();
        uint32_t v[4];
        uint32_t cnt = 0;
        for (int k = 0; k < 4; ++k) {
            v[k] = rowmap[lane * 4 + k];
            cnt += v[k];
        }
        uint32_t total;
        uint32_t ex = wave_excl_scan(cnt, &total);
        arena_sync<LDS>();
        for (int k = 0; k < 4; ++k) {
            rowmap[lane * 4 + k] = v[k] ? ex + 1 : 0;
            ex += v[k];
        }
        P.R = total + 1;
        P.map.rowmap = rowmap;
        pm = (uint64_t*)(arena + 1024);
    }
    else {
        const uint32_t H = 1u << hash_bits;
        uint64_t* keys = (uint64_t*)arena;
        uint32_t* rows = (uint32_t*)(keys + H);
        for (uint32_t i = lane; i < H; i += 64) keys[i] = EMPTY_KEY;
        arena_sync<LDS>();
        for (uint32_t i = lane; i < m; i += 64) {
            uint32_t c = (uint32_t)p[i];
            uint32_t mask = H - 1;
            uint32_t h = sym_hash(c, hash_bits);
            for (;;) {
                uint64_t prev = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)EMPTY_KEY,
                                          (unsigned long long)c);
                if (prev == EMPTY_KEY || prev == (uint64_t)c) break;
                h = (h + 1) & mask;
            }
        }
        arena_sync<LDS>();
        uint32_t base = 0;
        for (uint32_t s0 = 0; s0 < H; s0 += 64) {
            uint32_t i = s0 + lane;
            uint32_t f = (keys[i] != EMPTY_KEY) ? 1u : 0u;
            uint32_t total;
            uint32_t ex = wave_excl_scan(f, &total);
            if (f) rows[i] = base + ex + 1;
            base += total;
        }
        P.R = base + 1;
        P.map.keys = keys;
        P.map.rows = rows;
        P.map.bits = hash_bits;
        pm = (uint64_t*)(rows + H);
        pm = (uint64_t*)(((uintptr_t)pm + 7) & ~(uintptr_t)7);
    }
    arena_sync<LDS>();
    for (uint32_t i = lane; i < P.R * P.Wp; i += 64) pm[i] = 0;
    arena_sync<LDS>();
    for (uint32_t i = lane; i < m; i += 64) {
        uint32_t r = P.map.find((uint32_t)p[i]);
        atomicOr((unsigned long long*)&pm[(size_t)r * P.Wp + (i >> 6)], 1ull << (i & 63));
    }
    arena_sync<LDS>();
    P.pm = pm;
    return P;
}

// ---- per-item recurrences -------------------------------------------------

template <int MAXS, typename CharT, bool LDS>
__device__ uint64_t run_lev(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ t, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t VP[MAXS], VN[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) {
        VP[s] = ~0ull;
        VN[s] = 0;
    }
    const uint32_t lastw = P.W - 1;
    const uint32_t last_slot = lastw >> 6, last_lane = lastw & 63;
    const uint32_t lastbit = (m - 1) & 63;
    uint32_t score = m;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        const uint32_t cnt = min(64u, n - k0);
        uint64_t Xn[MAXS];
        {
            const uint64_t* pmrow = P.pm + (size_t)readlane(row, 0) * P.Wp;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmrow[s * 64 + lane] : 0;
        }
        for (uint32_t i = 0; i < cnt; ++i) {
            uint64_t Xc[MAXS];
            const uint64_t* pmnext = P.pm + (size_t)readlane(row, min(i + 1, 63u)) * P.Wp; // prefetch the next row
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                Xc[s] = Xn[s];
                Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmnext[s * 64 + lane] : 0;
            }
            uint64_t hn_c0 = 0, hp_c0 = 1;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < P.slots) {
                    const uint32_t w = s * 64 + lane;
                    const bool act = w < P.W;
                    const uint64_t X = Xc[s];
                    const uint64_t vp = VP[s], vn = VN[s];
                    const uint64_t X1 = X | 1;
                    const uint64_t d0a = (((X & vp) + vp) ^ vp) | X | vn;
                    const uint64_t d0b = (((X1 & vp) + vp) ^ vp) | X1 | vn;
                    const bool ga = act && ((d0a & vp) >> 63);
                    const bool gb = act && ((d0b & vp) >> 63);
                    const uint64_t G = ballot(ga), Pp = ballot(gb && !ga);
                    uint64_t sum;
                    const uint64_t cout = add3_carry(G | Pp, G, hn_c0, &sum);
                    const uint64_t cin = sum ^ Pp;
                    const uint32_t hin = lane_bit(cin, lane);
                    const uint64_t D0 = hin ? d0b : d0a;
                    uint64_t HP = vn | ~(D0 | vp);
                    uint64_t HN = D0 & vp;
                    const uint64_t HPm = ballot(act && (HP >> 63));
                    const uint32_t hpin = lane_bit((HPm << 1) | hp_c0, lane);
                    if ((uint32_t)s == last_slot && lane == last_lane) {
                        score += (uint32_t)(HP >> lastbit) & 1u;
                        score -= (uint32_t)(HN >> lastbit) & 1u;
                    }
                    HP = (HP << 1) | hpin;
                    HN = (HN << 1) | hin;
                    VP[s] = HN | ~(D0 | HP);
                    VN[s] = HP & D0;
                    hn_c0 = cout;
                    hp_c0 = HPm >> 63;
                }
            }
        }
    }
    return readlane(score, last_lane);
}

// Levenshtein with a cutoff k (|m - n| <= k < max(m, n)): exact when the
// distance is <= k, otherwise some value > k.  Ukkonen's band in the spirit of
// Levenshtein_impl.hpp:641-850, formulated for words across lanes:
//   * a cell (i, j) can lie on a path of cost <= k only if
//     |i - j| + |(m - n) - (i - j)| <= k, i.e. i - j in [dlo, dhi]; at text
//     column j only the words overlapping rows [j + dlo, j + dhi] are updated
//     (whole slots outside them are skipped);
//   * words above the band are frozen and the first active word takes the
//     horizontal carry-in (HP = 1, HN = 0); a word entering the band at the
//     bottom starts from the word above (VP = all ones).  Both only
//     over-estimate cells, so cells on a path of cost <= k stay exact;
//   * each word keeps the value of its bottom cell; a word whose cells are all
//     > k (bottom - rows + 1 > k) is dropped from the top of the band, and an
//     empty band ends the pair early with k + 1.
template <int MAXS, typename CharT, bool LDS>
__device__ uint64_t run_lev_band(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ t, uint32_t n,
                                 uint32_t k)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t W = P.W, lastw = W - 1, lastbit = (m - 1) & 63;
    const int32_t dm = (int32_t)m - (int32_t)n;
    const int32_t dlo = -(int32_t)((uint32_t)((int32_t)k - dm) >> 1); // ceil((dm - k) / 2)
    const int32_t dhi = (int32_t)((uint32_t)((int32_t)k + dm) >> 1);  // floor((dm + k) / 2)
    uint64_t VP[MAXS], VN[MAXS];
    uint32_t sc[MAXS]; // value of the word's bottom cell in the current column
#pragma unroll
    for (int s = 0; s < MAXS; ++s) {
        const uint32_t w = s * 64 + lane;
        VP[s] = ~0ull;
        VN[s] = 0;
        sc[s] = min((w + 1) * 64, m); // column 0: D[i][0] = i
    }
    auto word_lo = [&](uint32_t j) { return (uint32_t)(max(1, (int32_t)j + dlo) - 1) >> 6; };
    auto word_hi = [&](uint32_t j) { return (uint32_t)(min((int32_t)m, (int32_t)j + dhi) - 1) >> 6; };
    uint32_t fb = word_lo(1), lb = word_hi(1);
    // PM words of the next text unit are loaded one column ahead (slots of the
    // current band and the one below it, where the band can grow)
    auto load_row = [&](uint32_t r, uint64_t* Xd) {
        const uint64_t* pmrow = P.pm + (size_t)r * P.Wp;
        const uint32_t s0 = fb >> 6, s1 = (lb >> 6) + 1;
#pragma unroll
        for (int s = 0; s < MAXS; ++s) {
            const uint32_t w = s * 64 + lane;
            Xd[s] = ((uint32_t)s >= s0 && (uint32_t)s <= s1 && w < P.W) ? pmrow[w] : 0;
        }
    };
    uint64_t Xn[MAXS];
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        const uint32_t cnt = min(64u, n - k0);
        load_row(readlane(row, 0), Xn);
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = k0 + i + 1; // the column this text unit produces
            uint64_t Xc[MAXS];
#pragma unroll
            for (int s = 0; s < MAXS; ++s) Xc[s] = Xn[s];
            if (i + 1 < cnt) load_row(readlane(row, i + 1), Xn);
            fb = max(fb, word_lo(j));
            const uint32_t nl = word_hi(j);
            for (; lb < nl; ++lb) {
                // word lb + 1 enters the band: previous column from the word above
                const uint32_t wa = lb, wn = lb + 1;
                uint32_t v = 0;
#pragma unroll
                for (int s = 0; s < MAXS; ++s)
                    if ((uint32_t)s == (wa >> 6)) v = sc[s];
                const uint32_t above = readlane(v, wa & 63);
                const uint32_t rows = wn == lastw ? m - wn * 64 : 64u;
#pragma unroll
                for (int s = 0; s < MAXS; ++s) {
                    if ((uint32_t)s == (wn >> 6) && lane == (wn & 63)) {
                        VP[s] = ~0ull;
                        VN[s] = 0;
                        sc[s] = above + rows;
                    }
                }
            }
            if (fb > lb) return (uint64_t)k + 1;
            const uint32_t sf = fb >> 6, sl = lb >> 6;
            uint64_t hn_c0 = 0, hp_c0 = 1;
            uint64_t alive[MAXS];
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                alive[s] = 0;
                if ((uint32_t)s < sf || (uint32_t)s > sl) continue;
                const uint32_t w = s * 64 + lane;
                const bool act = w >= fb && w <= lb;
                const uint64_t X = act ? Xc[s] : 0;
                const uint64_t vp = VP[s], vn = VN[s];
                const uint64_t X1 = X | 1;
                const uint64_t d0a = (((X & vp) + vp) ^ vp) | X | vn;
                const uint64_t d0b = (((X1 & vp) + vp) ^ vp) | X1 | vn;
                const bool ga = act && ((d0a & vp) >> 63);
                const bool gb = act && ((d0b & vp) >> 63);
                const uint64_t G = ballot(ga), Pp = ballot(gb && !ga);
                uint64_t sum;
                const uint64_t cout = add3_carry(G | Pp, G, hn_c0, &sum);
                const uint64_t cin = sum ^ Pp;
                const uint32_t hin = lane_bit(cin, lane);
                const uint64_t D0 = hin ? d0b : d0a;
                uint64_t HP = vn | ~(D0 | vp);
                uint64_t HN = D0 & vp;
                const uint64_t HPm = ballot(act && (HP >> 63));
                const uint64_t first = (uint32_t)s == sf ? (1ull << (fb & 63)) : 0ull;
                const uint32_t hpin = lane_bit((HPm << 1) | hp_c0 | first, lane);
                const uint32_t hb = w == lastw ? lastbit : 63u;
                const int32_t dv = (int32_t)((HP >> hb) & 1) - (int32_t)((HN >> hb) & 1);
                HP = (HP << 1) | hpin;
                HN = (HN << 1) | hin;
                if (act) {
                    VP[s] = HN | ~(D0 | HP);
                    VN[s] = HP & D0;
                    sc[s] = (uint32_t)((int32_t)sc[s] + dv);
                }
                const uint32_t rows = w == lastw ? m - w * 64 : 64u;
                alive[s] = ballot(act && sc[s] + 1 <= k + rows); // some cell of the word is <= k
                hn_c0 = cout;
                hp_c0 = HPm >> 63;
            }
            // drop dead words from the top of the band
            bool found = false;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if (!found && (uint32_t)s >= sf && (uint32_t)s <= sl) {
                    const uint64_t a = alive[s] & ((uint32_t)s == sf ? (~0ull << (fb & 63)) : ~0ull);
                    if (a) {
                        fb = (uint32_t)s * 64 + (uint32_t)__builtin_ctzll(a);
                        found = true;
                    }
                }
            }
            if (!found) return (uint64_t)k + 1;
        }
    }
    if (lastw < fb || lastw > lb) return (uint64_t)k + 1;
    uint32_t v = 0;
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
        if ((uint32_t)s == (lastw >> 6)) v = sc[s];
    return readlane(v, lastw & 63);
}

template <int MAXS, typename CharT, bool LDS>
__device__ uint64_t run_lcs(const Pattern<CharT, LDS>& P, const CharT* __restrict__ t, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t S[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) S[s] = ~0ull;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        const uint32_t cnt = min(64u, n - k0);
        uint64_t Xn[MAXS];
        {
            const uint64_t* pmrow = P.pm + (size_t)readlane(row, 0) * P.Wp;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmrow[s * 64 + lane] : 0;
        }
        for (uint32_t i = 0; i < cnt; ++i) {
            uint64_t Xc[MAXS];
            const uint64_t* pmnext = P.pm + (size_t)readlane(row, min(i + 1, 63u)) * P.Wp; // prefetch the next row
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                Xc[s] = Xn[s];
                Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmnext[s * 64 + lane] : 0;
            }
            uint64_t c0 = 0;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < P.slots) {
                    const uint32_t w = s * 64 + lane;
                    const bool act = w < P.W;
                    const uint64_t Sv = S[s];
                    const uint64_t u = Sv & Xc[s];
                    uint64_t x = Sv + u;
                    const bool g = act && (x < Sv);
                    const bool p = act && !g && (x == ~0ull);
                    const uint64_t G = ballot(g), Pp = ballot(p);
                    uint64_t sum;
                    const uint64_t cout = add3_carry(G | Pp, G, c0, &sum);
                    x += lane_bit(sum ^ Pp, lane);
                    if (act) S[s] = x | (Sv - u);
                    c0 = cout;
                }
            }
        }
    }
    uint32_t cnt = 0;
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
        if ((uint32_t)s < P.slots && (uint32_t)(s * 64) + lane < P.W) cnt += __popcll(~S[s]);
    for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d, 64);
    return cnt;
}

template <int MAXS, typename CharT, bool LDS>
__device__ uint64_t run_osa(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ t, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t VP[MAXS], VN[MAXS], D0s[MAXS], PMo[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) {
        VP[s] = ~0ull;
        VN[s] = 0;
        D0s[s] = 0;
        PMo[s] = 0;
    }
    const uint32_t lastw = P.W - 1;
    const uint32_t last_slot = lastw >> 6, last_lane = lastw & 63;
    const uint32_t lastbit = (m - 1) & 63;
    uint32_t score = m;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        const uint32_t cnt = min(64u, n - k0);
        uint64_t Xn[MAXS];
        {
            const uint64_t* pmrow = P.pm + (size_t)readlane(row, 0) * P.Wp;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmrow[s * 64 + lane] : 0;
        }
        for (uint32_t i = 0; i < cnt; ++i) {
            uint64_t Xc[MAXS];
            const uint64_t* pmnext = P.pm + (size_t)readlane(row, min(i + 1, 63u)) * P.Wp; // prefetch the next row
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                Xc[s] = Xn[s];
                Xn[s] = ((uint32_t)s < P.slots && s * 64 + lane < P.W) ? pmnext[s * 64 + lane] : 0;
            }
            uint64_t hn_c0 = 0, hp_c0 = 1, tr_c0 = 0;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < P.slots) {
                    const uint32_t w = s * 64 + lane;
                    const bool act = w < P.W;
                    const uint64_t X = Xc[s];
                    const uint64_t vp = VP[s], vn = VN[s], d0old = D0s[s];
                    // ((~D0_last) & PM_last) >> 63 of the word below in the pattern
                    const uint64_t TRm = ballot(act && ((((~d0old) & X) >> 63) & 1));
                    const uint32_t trin = lane_bit((TRm << 1) | tr_c0, lane);
                    const uint64_t TR = ((((~d0old) & X) << 1) | trin) & PMo[s];
                    const uint64_t X1 = X | 1;
                    const uint64_t d0a = (((X & vp) + vp) ^ vp) | X | vn | TR;
                    const uint64_t d0b = (((X1 & vp) + vp) ^ vp) | X1 | vn | TR;
                    const bool ga = act && ((d0a & vp) >> 63);
                    const bool gb = act && ((d0b & vp) >> 63);
                    const uint64_t G = ballot(ga), Pp = ballot(gb && !ga);
                    uint64_t sum;
                    const uint64_t cout = add3_carry(G | Pp, G, hn_c0, &sum);
                    const uint32_t hin = lane_bit(sum ^ Pp, lane);
                    const uint64_t D0 = hin ? d0b : d0a;
                    uint64_t HP = vn | ~(D0 | vp);
                    uint64_t HN = D0 & vp;
                    const uint64_t HPm = ballot(act && (HP >> 63));
                    const uint32_t hpin = lane_bit((HPm << 1) | hp_c0, lane);
                    if ((uint32_t)s == last_slot && lane == last_lane) {
                        score += (uint32_t)(HP >> lastbit) & 1u;
                        score -= (uint32_t)(HN >> lastbit) & 1u;
                    }
                    HP = (HP << 1) | hpin;
                    HN = (HN << 1) | hin;
                    VP[s] = HN | ~(D0 | HP);
                    VN[s] = HP & D0;
                    D0s[s] = D0;
                    PMo[s] = X;
                    hn_c0 = cout;
                    hp_c0 = HPm >> 63;
                    tr_c0 = TRm >> 63;
                }
            }
        }
    }
    return readlane(score, last_lane);
}

// Jaro counts.  tflag: per-wave scratch of ceil(n/64) words.
template <int MAXS, typename CharT, bool LDS>
__device__ RawResult run_jaro(const Pattern<CharT, LDS>& P, uint32_t m, const CharT* __restrict__ t, uint32_t n,
                              uint64_t* tflag)
{
    const uint32_t lane = threadIdx.x & 63;
    RawResult res{0, 0, 0};
    if (m == 0 || n == 0) return res;
    uint32_t Bound = max(m, n) / 2;
    if (Bound > 0) Bound--;
    uint64_t Pf[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) Pf[s] = 0;
    const uint32_t maxpfx = min(min(m, n), 4u);
    uint32_t pfx = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        const uint32_t cnt = min(64u, n - k0);
        uint64_t tacc = 0;
        for (uint32_t i = 0; i < cnt; ++i) {
            const uint32_t j = k0 + i;
            const uint32_t r = readlane(row, i);
            const uint64_t* pmrow = P.pm + (size_t)r * P.Wp;
            if (j < maxpfx) {
                uint64_t w0 = pmrow[0];
                if (pfx == j && ((w0 >> j) & 1)) pfx = j + 1;
            }
            // window [lo, hi] over pattern positions
            const int64_t lo = (int64_t)j - (int64_t)Bound;
            const int64_t hi = (int64_t)j + (int64_t)Bound;
            bool found = false;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < P.slots && !found) {
                    const uint32_t w = s * 64 + lane;
                    const int64_t base = (int64_t)w * 64;
                    const int64_t a = lo - base, b = hi - base;
                    uint64_t wm = 0;
                    if (b >= 0 && a <= 63) {
                        const uint32_t aa = a < 0 ? 0 : (uint32_t)a;
                        const uint32_t bb = b > 63 ? 63 : (uint32_t)b;
                        wm = (~0ull >> (63 - bb)) & (~0ull << aa);
                    }
                    const uint64_t x = (w < P.W ? pmrow[w] : 0) & wm & ~Pf[s];
                    const uint64_t any = ballot(w < P.W && x != 0);
                    if (any) {
                        const uint32_t first = __builtin_ctzll(any);
                        if (lane == first) Pf[s] |= blsi(x);
                        found = true;
                    }
                }
            }
            if (found) tacc |= 1ull << i;
        }
        if (lane == 0) tflag[k0 >> 6] = tacc;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    // transposition walk: k-th flagged text position against k-th flagged pattern position
    uint64_t Pr[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) Pr[s] = Pf[s];
    uint32_t tr = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint64_t tw = ((volatile uint64_t*)tflag)[k0 >> 6];
        if (!tw) continue;
        uint32_t row = 0;
        if (k0 + lane < n) row = P.map.find((uint32_t)t[k0 + lane]);
        while (tw) {
            const uint32_t i = __builtin_ctzll(tw);
            tw &= tw - 1;
            const uint32_t r = readlane(row, i);
            const uint64_t* pmrow = P.pm + (size_t)r * P.Wp;
            bool done = false;
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < P.slots && !done) {
                    const uint64_t any = ballot(Pr[s] != 0);
                    if (any) {
                        const uint32_t first = __builtin_ctzll(any);
                        uint32_t mis = 0;
                        if (lane == first) {
                            const uint64_t pmask = blsi(Pr[s]);
                            mis = ((s * 64 + lane < P.W ? pmrow[s * 64 + lane] : 0) & pmask) == 0;
                            Pr[s] ^= pmask;
                        }
                        tr += readlane(mis, first);
                        done = true;
                    }
                }
            }
        }
    }
    uint32_t cc = 0;
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
        if ((uint32_t)s < P.slots) cc += __popcll(Pf[s]);
    for (int d = 32; d >= 1; d >>= 1) cc += __shfl_xor(cc, d, 64);
    res.v = cc;
    res.tr = tr;
    res.prefix = pfx;
    return res;
}

// Length filter: when the answer depends on the raw distance only through
// min(d, K+1) and |m - n| > K, the pair needs no recurrence at all
// (the reference's own early exits: Levenshtein_impl.hpp:863, 910;
// LCSseq_impl.hpp:382-389).  Returns true and the raw value to use.
__device__ __forceinline__ bool length_filter(const ScoreSpec& s, int algo, uint64_t la, uint64_t lb, RawResult* out)
{
    if (s.semantics != RFA_CACHED || algo == ALGO_JARO) return false;
    const uint64_t diff = la > lb ? la - lb : lb - la;
    uint64_t K = s.band_k; // bound on the raw distance, computed on the host for fixed-cutoff kinds
    if (s.kind == RFA_NORMALIZED_DISTANCE || s.kind == RFA_NORMALIZED_SIMILARITY || s.metric == RFA_RATIO) {
        // cutoff in units of the maximum
        ScoreSpec t = s;
        if (s.metric == RFA_RATIO) t.metric = RFA_INDEL;
        IntPair p{&t, la, lb, 0};
        uint64_t maximum = int_maximum(p);
        double c = s.cutoff_d;
        if (s.metric == RFA_RATIO) c = s.cutoff_d / 100;
        if (s.kind == RFA_NORMALIZED_SIMILARITY || s.metric == RFA_RATIO) c = normsim_to_normdist(c);
        K = (uint64_t)ceil((double)maximum * c);
        if (s.metric == RFA_LEVENSHTEIN) {
            if (s.ins == 0) return false;
            K = ceil_div_u64(K, s.ins);
        }
    }
    else if (s.kind == RFA_SIMILARITY && s.metric != RFA_LCSSEQ) {
        IntPair p{&s, la, lb, 0};
        uint64_t maximum = int_maximum(p);
        if (s.cutoff_u > maximum) {
            out->v = algo == ALGO_LCS ? 0 : la + lb; // result is 0 whatever the value
            return true;
        }
        K = maximum - s.cutoff_u;
        if (s.metric == RFA_LEVENSHTEIN) {
            if (s.ins == 0) return false;
            K = ceil_div_u64(K, s.ins);
        }
    }
    if (algo == ALGO_LCS) {
        if (s.metric == RFA_LCSSEQ) {
            // similarity: LCS <= min(m, n); distance: max - LCS >= diff
            return false;
        }
        // Indel distance = m + n - 2 LCS >= diff; indel-weighted Levenshtein likewise
        if (diff > K) {
            out->v = 0; // distance m + n > K: capped identically
            return true;
        }
        return false;
    }
    if (algo == ALGO_LEV || algo == ALGO_OSA) {
        if (diff > K) {
            out->v = K + 1;
            return true;
        }
    }
    return false;
}

// Arena bytes one pattern of m units needs: symbol map + PM rows (rows bounded by
// the distinct-symbol count, which is at most min(m, alphabet)).
template <typename CharT>
__host__ __device__ inline uint32_t pattern_hash_bits(uint32_t m)
{
    const uint64_t alpha = sizeof(CharT) == 1 ? 256 : (sizeof(CharT) == 2 ? 65536 : 0xffffffffull);
    const uint64_t d = m < alpha ? m : alpha;
    uint32_t b = 6;
    while ((1ull << b) < 2 * d) ++b;
    return b;
}

template <typename CharT>
__host__ __device__ inline uint64_t pattern_arena_bytes(uint32_t m)
{
    const uint64_t alpha = sizeof(CharT) == 1 ? 256 : (sizeof(CharT) == 2 ? 65536 : 0xffffffffull);
    const uint64_t rows = (m < alpha ? m : alpha) + 1;
    const uint64_t W = (m + 63) / 64;
    const uint64_t map = sizeof(CharT) == 1 ? 1024 : ((1ull << pattern_hash_bits<CharT>(m)) * 12 + 8);
    return map + rows * W * 8;
}

// Arena bytes of an 8-bit pattern from its distinct symbols (rows = distinct + 1):
// the row map, then one PM row per symbol.  The whole wave calls it.
__device__ __forceinline__ uint64_t u8_arena_need(const uint8_t* __restrict__ p, uint32_t m, uint32_t* present)
{
    const uint32_t lane = threadIdx.x & 63;
    if (lane < 8) present[lane] = 0;
    __syncthreads();
    for (uint32_t i = lane; i < m; i += 64) {
        const uint32_t c = p[i];
        atomicOr(&present[c >> 5], 1u << (c & 31));
    }
    __syncthreads();
    uint32_t cnt = lane < 8 ? (uint32_t)__popc(present[lane]) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    return 1024 + (uint64_t)(cnt + 1) * ((m + 63) / 64) * 8;
}

// Q-gram count filter for the banded recurrence (Ukkonen, "Approximate string
// matching with q-grams and maximal matches", 1992): if ed(p, t) <= k, every
// edit touches at most Q of t's Q-grams, so at least (n - Q + 1) - k*Q of them
// occur verbatim in p.  The pattern's Q-grams are hashed into a 2^16-bit set;
// a hash collision can only add hits, so a pair with fewer hits than the bound
// has distance > k for certain, and its answer is k + 1 without the recurrence.
constexpr uint32_t QGRAM_Q = 8;
constexpr uint32_t QGRAM_WORDS = 2048; // 65,536 bits

template <typename CharT>
__device__ __forceinline__ uint32_t qgram_hash(const CharT* s)
{
    uint32_t u[QGRAM_Q];
    load_units<CharT, QGRAM_Q>(s, u);
    uint64_t x = 0;
#pragma unroll
    for (uint32_t i = 0; i < QGRAM_Q; ++i) x = sizeof(CharT) == 1 ? (x | ((uint64_t)u[i] << (8 * i))) : (x * 0x100000001B3ull) ^ u[i];
    return (uint32_t)((x * 0x9E3779B97F4A7C15ull) >> 48);
}

// True when the filter proves ed(p, t) > k.  The whole wave calls it; `bits`
// caches the set of the pattern named by `key` (bkey).
template <typename CharT>
__device__ bool qgram_rules_out(const CharT* __restrict__ p, uint32_t m, const CharT* __restrict__ t, uint32_t n,
                                uint64_t k, uint32_t* bits, uint64_t key, uint64_t& bkey)
{
    if (m < QGRAM_Q || n < QGRAM_Q) return false;
    const int64_t need = (int64_t)(n - QGRAM_Q + 1) - (int64_t)k * QGRAM_Q;
    if (need <= 0) return false;
    const uint32_t lane = threadIdx.x & 63;
    if (key != bkey) {
        for (uint32_t i = lane; i < QGRAM_WORDS; i += 64) bits[i] = 0;
        __syncthreads();
        for (uint32_t i = lane; i + QGRAM_Q <= m; i += 64) {
            const uint32_t h = qgram_hash(p + i);
            atomicOr(&bits[h >> 5], 1u << (h & 31));
        }
        __syncthreads();
        bkey = key;
    }
    uint32_t hits = 0;
    for (uint32_t i = lane; i + QGRAM_Q <= n; i += 64) {
        const uint32_t h = qgram_hash(t + i);
        hits += (bits[h >> 5] >> (h & 31)) & 1u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off, 64);
    return (int64_t)uniform(hits) < need;
}

// One item with the arena in LDS (LDS = true) or in the workgroup's global slice.
template <int ALGO, int MAXS, typename CharT, bool LDS, bool BAND, bool QG>
__device__ __forceinline__ RawResult block_item(int algo, const CharT* pp, uint32_t m, const CharT* tt, uint32_t n,
                                                uint8_t* arena, uint64_t* tflag, uint64_t key, bool reuse,
                                                Pattern<CharT, LDS>& P, uint64_t& prev_key, uint64_t band,
                                                uint32_t* qbits, uint64_t& qkey)
{
    RawResult res{0, 0, 0};
    if (key != prev_key || !reuse) {
        P = build_pattern<CharT, LDS>(pp, m, arena, pattern_hash_bits<CharT>(m));
        prev_key = key;
    }
    // Levenshtein needed exactly only up to `band`: the banded recurrence
    auto lev = [&]() -> uint64_t {
        if constexpr (BAND) {
            if (band < (uint64_t)max(m, n)) {
                if ((uint64_t)(m > n ? m - n : n - m) > band) return band + 1;
                if constexpr (QG) {
                    if (qgram_rules_out(pp, m, tt, n, band, qbits, key, qkey)) return band + 1;
                }
                return run_lev_band<MAXS>(P, m, tt, n, (uint32_t)band);
            }
        }
        return run_lev<MAXS>(P, m, tt, n);
    };
    if constexpr (ALGO == ALGO_MIXED) {
        if (algo == ALGO_LEV)
            res.v = lev();
        else if (algo == ALGO_OSA)
            res.v = run_osa<MAXS>(P, m, tt, n);
        else
            res.v = run_lcs<MAXS>(P, tt, n);
    }
    else if constexpr (ALGO == ALGO_LEV)
        res.v = lev();
    else if constexpr (ALGO == ALGO_LCS)
        res.v = run_lcs<MAXS>(P, tt, n);
    else if constexpr (ALGO == ALGO_OSA)
        res.v = run_osa<MAXS>(P, m, tt, n);
    else
        res = run_jaro<MAXS>(P, m, tt, n, tflag);
    return res;
}

// BAND: Levenshtein items with a cutoff below max(m, n) run the banded
// recurrence (a separate instantiation, so unbanded launches keep their
// register footprint).  QG: banded launches whose cutoffs are small enough for
// the q-gram filter to reject pairs (7k + 7 < m for some pattern) also hold its
// 8 KiB set in LDS.
template <int ALGO, int MAXS, typename CharT, bool BAND, bool QG>
__global__ __launch_bounds__(64) void k_block(BlockArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    uint8_t* lds_arena = (uint8_t*)smem_u64;
    uint8_t* glb_arena = a.pm_scratch ? (uint8_t*)(a.pm_scratch + (size_t)blockIdx.x * a.pm_scratch_words) : nullptr;
    uint64_t* tflag = a.tflag_scratch ? a.tflag_scratch + (size_t)blockIdx.x * a.tflag_words : nullptr;
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    __shared__ uint32_t s_item;
    __shared__ uint32_t present[8];
    __shared__ uint32_t qbits[QG ? QGRAM_WORDS : 1]; // q-gram set of the current pattern
    uint64_t keyL = ~0ull, keyG = ~0ull, need_key = ~0ull, need_u8 = 0, qkey = ~0ull;
    Pattern<CharT, true> PL{};
    Pattern<CharT, false> PG{};
    // items are taken from the queue a.batch at a time: a device-wide atomic per
    // item would cost more than the many items the length filter settles at once
    uint32_t next = 0, stop = 0;
    for (;;) {
        if (next >= stop) {
            if (lane == 0) s_item = atomicAdd(a.queue, a.batch);
            __syncthreads();
            next = uniform(s_item);
            __syncthreads();
            stop = next + a.batch;
        }
        const uint32_t item = next++;
        const uint32_t nitems = a.nitems_dev ? uniform(*a.nitems_dev) : a.nitems;
        if (item >= nitems) break;
        uint32_t ia, ib;
        uint64_t oidx;
        bool swap = false;
        if (a.item_packed) {
            const uint32_t pk = uniform(a.item_packed[item]);
            ia = ib = pk & 0x7fffffffu;
            swap = (pk >> 31) != 0;
            oidx = ia;
        }
        else if (a.cross_nb) {
            ia = a.item_a[item / a.cross_nb];
            ib = a.item_b[item % a.cross_nb];
            oidx = (uint64_t)ia * a.out_stride + ib;
        }
        else {
            ia = a.item_a[item];
            ib = a.item_b[item];
            oidx = a.out_index[item];
            swap = a.item_swap ? a.item_swap[item] != 0 : false;
        }
        ScoreSpec spec = a.spec;
        int algo = ALGO;
        if constexpr (ALGO == ALGO_MIXED) {
            spec.metric = uniform((uint32_t)a.pair_metric[ia]);
            algo = spec.metric == RFA_LEVENSHTEIN ? ALGO_LEV : (spec.metric == RFA_OSA ? ALGO_OSA : ALGO_LCS);
            if (spec.kind == RFA_DISTANCE && spec.metric != RFA_LCSSEQ) spec.band_k = spec.cutoff_u;
        }
        if (a.qcut) {
            spec.cutoff_u = a.qcut[ia];
            spec.band_k = band_from_cutoff(spec);
        }
        if (a.pcut) spec.cutoff_d = a.pcut[oidx]; // the LCS kinds use it only in the filter and epilogue
        const uint64_t oa = a.aoff[ia], ob = a.boff[ib];
        const uint32_t la = (uint32_t)(a.aoff[ia + 1] - oa), lb = (uint32_t)(a.boff[ib + 1] - ob);
        const CharT* pp = swap ? B + ob : A + oa;
        const CharT* tt = swap ? A + oa : B + ob;
        const uint32_t m = swap ? lb : la, n = swap ? la : lb;
        RawResult res{0, 0, 0};
        if (!length_filter(spec, algo, la, lb, &res)) {
            if (m == 0) {
                res.v = (algo == ALGO_LCS || algo == ALGO_JARO) ? 0 : n;
            }
            else {
                const uint64_t key = ((uint64_t)swap << 63) | (swap ? ib : ia);
                uint64_t need;
                if constexpr (sizeof(CharT) == 1) {
                    // 8-bit patterns are sized by their distinct symbols
                    if (key != need_key) {
                        need_u8 = u8_arena_need((const uint8_t*)pp, m, present);
                        need_key = key;
                    }
                    need = need_u8;
                }
                else {
                    need = pattern_arena_bytes<CharT>(m);
                }
                if (need <= a.lds_bytes)
                    res = block_item<ALGO, MAXS, CharT, true, BAND, QG>(algo, pp, m, tt, n, lds_arena, tflag, key,
                                                              a.reuse_pattern, PL, keyL, spec.band_k, qbits, qkey);
                else
                    res = block_item<ALGO, MAXS, CharT, false, BAND, QG>(algo, pp, m, tt, n, glb_arena, tflag, key,
                                                               a.reuse_pattern, PG, keyG, spec.band_k, qbits, qkey);
            }
        }
        if (lane == 0) finalize(spec, la, lb, res, a.out, oidx);
    }
}

template <int ALGO, int MAXS, typename CharT>
static void launch3(const BlockArgs& a, uint32_t nblocks, hipStream_t s)
{
    // banded Levenshtein: Levenshtein / mixed launches with a distance cutoff, patterns <= 32768 units
    if constexpr ((ALGO == ALGO_LEV || ALGO == ALGO_MIXED) && MAXS <= 8) {
        if (a.spec.kind == RFA_DISTANCE && a.spec.cutoff_u < (1ull << 31)) {
            if (a.qgram)
                hipLaunchKernelGGL((k_block<ALGO, MAXS, CharT, true, true>), dim3(nblocks), dim3(64), a.lds_bytes, s, a);
            else
                hipLaunchKernelGGL((k_block<ALGO, MAXS, CharT, true, false>), dim3(nblocks), dim3(64), a.lds_bytes, s,
                                   a);
            return;
        }
    }
    hipLaunchKernelGGL((k_block<ALGO, MAXS, CharT, false, false>), dim3(nblocks), dim3(64), a.lds_bytes, s, a);
}

template <int ALGO, typename CharT>
static void launch2(const BlockArgs& a, uint32_t nblocks, uint32_t slots, hipStream_t s)
{
    if (slots <= 1)
        launch3<ALGO, 1, CharT>(a, nblocks, s);
    else if (slots <= 2)
        launch3<ALGO, 2, CharT>(a, nblocks, s);
    else if (slots <= 4)
        launch3<ALGO, 4, CharT>(a, nblocks, s);
    else if (slots <= 8)
        launch3<ALGO, 8, CharT>(a, nblocks, s);
    else if (slots <= 16)
        launch3<ALGO, 16, CharT>(a, nblocks, s);
    else
        launch3<ALGO, 32, CharT>(a, nblocks, s);
}

// Instantiated once per code-unit width (rfa_block_u8/u16/u32.hip).
template <typename CharT>
void launch_block_w(int algo, const BlockArgs& a, uint32_t nblocks, uint32_t slots, hipStream_t s)
{
    switch (algo) {
    case ALGO_LEV: launch2<ALGO_LEV, CharT>(a, nblocks, slots, s); break;
    case ALGO_LCS: launch2<ALGO_LCS, CharT>(a, nblocks, slots, s); break;
    case ALGO_OSA: launch2<ALGO_OSA, CharT>(a, nblocks, slots, s); break;
    case ALGO_MIXED: launch2<ALGO_MIXED, CharT>(a, nblocks, slots, s); break;
    default: launch2<ALGO_JARO, CharT>(a, nblocks, slots, s); break;
    }
}

// rows: distinct symbols + 1 of the largest 8-bit pattern when known (0: the alphabet bound)
inline uint64_t block_arena_need(int char_width, uint32_t m, uint32_t rows = 0)
{
    if (char_width == 1 && rows) return 1024 + (uint64_t)rows * ((m + 63) / 64) * 8;
    if (char_width == 1) return pattern_arena_bytes<uint8_t>(m);
    if (char_width == 2) return pattern_arena_bytes<uint16_t>(m);
    return pattern_arena_bytes<uint32_t>(m);
}

} // namespace rfa
