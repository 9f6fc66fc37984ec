// rfa_band.hip -- banded Levenshtein distance of long queries (> 64 units,
// 8-bit units) against a choice set under an integer cutoff k: the calls
// CachedLevenshtein<uint8_t>(q).distance(s2, k) of a cdist with a cutoff
// (Levenshtein_impl.hpp:641-850 is the reference's banded block recurrence,
// :419-505 its one-word diagonal band, which this kernel widens to many words).
//
// Band.  A pair (query P of m units, choice T of n units, d = m - n, |d| <= k)
// reaches a distance <= k only through cells on the diagonals i - j in
// [dlo, dhi] = [ceil((d - k) / 2), floor((d + k) / 2)]: a path pays |i - j| to
// reach cell (i, j) and |d - (i - j)| to leave it.
//
// Diagonal window.  The pair keeps B = ceil((k + 2) / 64) 64-bit words that
// slide one row down per text unit: before text unit j (0-based) bit b of the
// window (word b / 64 in lane b / 64 of the pair's lanes) holds the vertical
// delta D[r][j] - D[r - 1][j] of row r = j + dlo + b.  Text unit j evaluates
// column j + 1 on the same rows (the reference's one-word band step: D0 from
// the addition over the window, HP and HN unshifted, D0 shifted down one row
// for the next window), with the addition's carry passed across the words.
// The window's top row takes no carry from above and the row entering at the
// bottom takes no diagonal zero; both only over-estimate cells on diagonals
// dlo - 1 and dlo + 64 B - 1, outside the band, so every cell on a path of
// cost <= k stays exact.  Rows at or above row 0 start with vertical delta 0
// and never match, so they copy row 0 (D[0][j] = j).
//
// Score and early exit.  D[m][n] lies on diagonal d; the lane holding that
// diagonal adds 1 - D0 of it per text unit.  Every lane also follows the
// bottom cell of its word (diagonal dL).  A word's cells on diagonals
// dL - 63 .. dL are at least bottom - (dL - diagonal) and still owe
// |d - diagonal| to reach (m, n); along any path that sum never decreases, so
// once it exceeds k for every word the distance is > k and the pair ends with
// k + 1 (checked every 8 text units).
//
// Layout.  One workgroup per (query, chunk of choices).  The query's PM rows
// live in LDS as dwords, padded with zero words on both sides so any window
// position reads consecutive dwords and two funnel shifts per word
// (v_alignbit_b32); a byte -> row-offset map selects the row.  Each lane holds
// WPL consecutive words (1 or 2), so a wave runs floor(64 / ceil(B / WPL))
// pairs at once: the addition carries between a pair's lanes resolve by ballot
// carry-lookahead, and the bit D0 shifts down from the next lane comes by DPP.  A pair that ends hands its lanes
// to the next candidate choice of the chunk (choices are length-filtered 64 at
// a time), so early exits and uneven lengths keep the wave busy.
#include "rfa_block_impl.hpp"
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"

namespace rfa {

// One text unit for a lane's WPL consecutive window words: their match bits are
// 2 WPL + 1 dwords of the symbol's PM row from bit offset g (two funnel shifts per
// word), then the band step over the lane's words with the addition carried through
// them (one chain of v_add_co / v_addc whose carry-out is the lane's generate mask).
// `upd` false leaves the pair's state as it is (its text ended earlier in the chunk).
template <bool MULTI, bool PRED, int WPL>
__device__ __forceinline__ void band_step(const uint8_t* __restrict__ lds, uint32_t row, uint32_t g, uint64_t LASTM,
                                          uint32_t lane, bool last, uint64_t (&VP)[WPL], uint64_t (&VN)[WPL],
                                          int32_t (&sc)[WPL], bool upd)
{
    // row: byte address of the symbol's PM row in the workgroup's LDS
    const uint32_t* __restrict__ p = (const uint32_t*)(lds + row + ((g >> 5) << 2));
    uint32_t w[2 * WPL + 1];
#pragma unroll
    for (int t = 0; t < 2 * WPL + 1; ++t) w[t] = p[t];
    uint32_t Xl[WPL], Xh[WPL], Pl[WPL], Ph[WPL], Nl[WPL], Nh[WPL], Al[WPL], Ah[WPL], sl[WPL], sh[WPL];
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        Xl[i] = __builtin_amdgcn_alignbit(w[2 * i + 1], w[2 * i], g);
        Xh[i] = __builtin_amdgcn_alignbit(w[2 * i + 2], w[2 * i + 1], g);
        Pl[i] = (uint32_t)VP[i];
        Ph[i] = (uint32_t)(VP[i] >> 32);
        Nl[i] = (uint32_t)VN[i];
        Nh[i] = (uint32_t)(VN[i] >> 32);
        Al[i] = Xl[i] & Pl[i];
        Ah[i] = Xh[i] & Ph[i];
    }
    if constexpr (MULTI || WPL > 1) {
        // (X & VP) + VP over the lane's words; then the carry into the lane, generated where the
        // lane's own add overflows and propagated where its sum is all ones (no overflow is
        // possible then), added in as the carry-in of one more chain
        uint64_t ga, t;
        if constexpr (WPL == 1) {
            asm("v_add_co_u32_e64 %0, %2, %3, %4\n\tv_addc_co_u32_e64 %1, %2, %5, %6, %2"
                : "=&v"(sl[0]), "=&v"(sh[0]), "=&s"(ga)
                : "v"(Al[0]), "v"(Pl[0]), "v"(Ah[0]), "v"(Ph[0]));
        }
        else {
            asm("v_add_co_u32_e64 %0, %4, %5, %6\n\tv_addc_co_u32_e64 %1, %4, %7, %8, %4\n\t"
                "v_addc_co_u32_e64 %2, %4, %9, %10, %4\n\tv_addc_co_u32_e64 %3, %4, %11, %12, %4"
                : "=&v"(sl[0]), "=&v"(sh[0]), "=&v"(sl[1]), "=&v"(sh[1]), "=&s"(ga)
                : "v"(Al[0]), "v"(Pl[0]), "v"(Ah[0]), "v"(Ph[0]), "v"(Al[1]), "v"(Pl[1]), "v"(Ah[1]), "v"(Ph[1]));
        }
        if constexpr (MULTI) {
            bool ones = true;
#pragma unroll
            for (int i = 0; i < WPL; ++i) ones = ones && (sl[i] & sh[i]) == ~0u;
            const uint64_t G = ga & ~LASTM, Pp = ballot(ones) & ~LASTM;
            const uint64_t cin = ((G | Pp) + G) ^ Pp;
            uint32_t fl[WPL], fh[WPL];
            if constexpr (WPL == 1) {
                asm("v_addc_co_u32_e64 %0, %2, %3, 0, %4\n\tv_addc_co_u32_e64 %1, %2, %5, 0, %2"
                    : "=&v"(fl[0]), "=&v"(fh[0]), "=&s"(t)
                    : "v"(sl[0]), "s"(cin), "v"(sh[0]));
            }
            else {
                asm("v_addc_co_u32_e64 %0, %4, %5, 0, %6\n\tv_addc_co_u32_e64 %1, %4, %7, 0, %4\n\t"
                    "v_addc_co_u32_e64 %2, %4, %8, 0, %4\n\tv_addc_co_u32_e64 %3, %4, %9, 0, %4"
                    : "=&v"(fl[0]), "=&v"(fh[0]), "=&v"(fl[1]), "=&v"(fh[1]), "=&s"(t)
                    : "v"(sl[0]), "s"(cin), "v"(sh[0]), "v"(sl[1]), "v"(sh[1]));
            }
#pragma unroll
            for (int i = 0; i < WPL; ++i) {
                sl[i] = fl[i];
                sh[i] = fh[i];
            }
        }
    }
    else {
        const uint64_t sum = ((((uint64_t)Ah[0]) << 32) | Al[0]) + VP[0];
        sl[0] = (uint32_t)sum;
        sh[0] = (uint32_t)(sum >> 32);
    }
    uint32_t D0l[WPL], D0h[WPL];
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        D0l[i] = bop3<T_XOR_OR>(sl[i], Pl[i], Xl[i] | Nl[i]);
        D0h[i] = bop3<T_XOR_OR>(sh[i], Ph[i], Xh[i] | Nh[i]);
    }
    // D0 one row down: bit 0 of the next word enters at the top of each word -- the
    // lane's next word, or the next lane's first word (DPP wave_shl:1); the last word of
    // a pair takes none
    uint32_t in = 0;
    if constexpr (MULTI) {
        const uint32_t nb = (uint32_t)__builtin_amdgcn_mov_dpp((int)D0l[0], 0x130, 0xF, 0xF, true); // lane 63: 0
        in = last ? 0u : nb;
    }
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        const uint32_t HPl = bop3<T_OR_NOR>(Nl[i], D0l[i], Pl[i]), HPh = bop3<T_OR_NOR>(Nh[i], D0h[i], Ph[i]);
        const uint32_t HNl = D0l[i] & Pl[i], HNh = D0h[i] & Ph[i];
        const uint32_t next = i + 1 < WPL ? D0l[i + 1 < WPL ? i + 1 : 0] : in;
        const uint32_t Sl = __builtin_amdgcn_alignbit(D0h[i], D0l[i], 1), Sh = __builtin_amdgcn_alignbit(next, D0h[i], 1);
        const uint64_t nVP = ((uint64_t)bop3<T_OR_NOR>(HNh, Sh, HPh) << 32) | bop3<T_OR_NOR>(HNl, Sl, HPl);
        const uint64_t nVN = ((uint64_t)(Sh & HPh) << 32) | (Sl & HPl);
        const int32_t nsc = sc[i] + 1 + ((int32_t)D0h[i] >> 31);
        if constexpr (PRED) {
            VP[i] = upd ? nVP : VP[i];
            VN[i] = upd ? nVN : VN[i];
            sc[i] = upd ? nsc : sc[i];
        }
        else {
            VP[i] = nVP;
            VN[i] = nVN;
            sc[i] = nsc;
        }
    }
}

// The diagonal window above is worth its bookkeeping while it stays narrow:
// band_words(k) <= 64 words, WPL consecutive words per lane.
template <bool MULTI, int WPL>
__device__ __forceinline__ void band_run(const BandArgs& a, uint32_t* smem_band)
{
    uint32_t* rowmap = smem_band;    // [256] LDS byte address of each symbol's PM row (row 0: zeros)
    uint32_t* pmd = smem_band + 256; // PM rows, 2 * WP dwords each
    __shared__ uint32_t s_next, s_rows;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t qslot = blockIdx.x / a.chunks_per_query;
    const uint32_t chunk = blockIdx.x - qslot * a.chunks_per_query;
    const uint32_t qi = a.queries[qslot];
    const uint64_t qo = a.qoff[qi];
    const uint32_t m = (uint32_t)(a.qoff[qi + 1] - qo);
    const uint8_t* __restrict__ P = a.qchars + qo;
    ScoreSpec spec = a.spec;
    if (a.qcut) {
        spec.cutoff_u = a.qcut[qi];
        spec.band_k = band_from_cutoff(spec);
    }
    const uint32_t k = (uint32_t)spec.band_k; // the host launches queries with band_words(k) <= 64 only
    const uint32_t B = band_words(k), PAD = band_pad(k), WP = band_row_words(m, k, WPL);
    const uint32_t c0 = chunk * a.chunk, c1 = min(a.nc, c0 + a.chunk);
    uint64_t* __restrict__ out = (uint64_t*)a.out;

    // ---- the query's PM rows ----
    for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) rowmap[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) rowmap[P[i]] = 1;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t v[4], cnt = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = rowmap[lane * 4 + q];
            cnt += v[q];
        }
        uint32_t total;
        uint32_t ex = wave_excl_scan(cnt, &total);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rowmap[lane * 4 + q] = 1024u + 4u * (v[q] ? (ex + 1) * 2 * WP : 0u);
            ex += v[q];
        }
        if (lane == 0) {
            s_rows = total + 1;
            s_next = c0;
        }
    }
    __syncthreads();
    const uint32_t ndw = s_rows * 2 * WP;
    for (uint32_t i = threadIdx.x; i < ndw; i += blockDim.x) pmd[i] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const uint32_t bit = 64 * PAD + i;
        atomicOr(&pmd[(rowmap[P[i]] - 1024u) / 4 + (bit >> 5)], 1u << (bit & 31));
    }
    __syncthreads();

    // ---- pairs of S = ceil(B / WPL) lanes (a window of S * WPL words) ----
    const uint32_t S = (B + WPL - 1) / WPL, nseg = 64 / S;
    const uint32_t seg = lane / S, L = lane - seg * S;
    const bool lane_ok = seg < nseg;
    const uint64_t LASTM = ballot(lane_ok && L == S - 1);
    const bool last = !lane_ok || L == S - 1;
    const uint32_t seg_lane0 = lane_ok ? seg * S : lane;
    const uint64_t seg_mask = S == 64 ? ~0ull : ((1ull << S) - 1);
    const uint64_t lt_mask = (1ull << lane) - 1;

    bool busy = false;
    uint32_t ci = 0, n = 0, jt = 0, iD = 0;
    bool isD = false;
    const uint8_t* T = P;
    int32_t g0 = 0, sc[WPL], scmax[WPL];
    uint64_t dmask = 0; // bits of word iD between diagonal d and the word's bottom
    uint64_t VP[WPL], VN[WPL];
#pragma unroll
    for (int i = 0; i < WPL; ++i) {
        sc[i] = scmax[i] = 0;
        VP[i] = VN[i] = 0;
    }
    // candidates of the chunk: lane r holds the r-th pending one (wave-uniform counts)
    uint32_t pend_ci = 0, pend_n = 0, pend_off = 0;
    bool exhausted = false;
    uint64_t cnt_steps = 0, cnt_pairs = 0; // RFA_OPTION_COUNT_WORD_STEPS

    for (;;) {
        // ---- hand idle lanes the next candidates ----
        for (;;) {
            const uint64_t idle = ballot(lane_ok && L == 0 && !busy);
            if (idle == 0) break;
            if (pend_off >= pend_n) {
                if (exhausted) break;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&s_next, 64u);
                base = readlane(base, 0);
                if (base >= c1) {
                    exhausted = true;
                    break;
                }
                // 64 choices: the lengths settle most of them
                const uint32_t cj = base + lane;
                bool cand = false;
                if (cj < c1) {
                    const uint64_t o = a.coff[cj];
                    const uint32_t nn = (uint32_t)(a.coff[cj + 1] - o);
                    // the reference's length filter (Levenshtein_impl.hpp:921-922); an
                    // empty choice is m deletions
                    const uint32_t diff = m > nn ? m - nn : nn - m;
                    const bool decided = diff > k || nn == 0;
                    if (decided) out[(uint64_t)qi * a.out_stride + cj] = finalize_int(spec, m, nn, diff > k ? (uint64_t)k + 1 : m);
                    cand = !decided;
                }
                const uint64_t cm = ballot(cand);
                const uint32_t np = (uint32_t)__popcll(cm);
                const uint32_t dst = cand ? (uint32_t)__popcll(cm & lt_mask) : np + (uint32_t)__popcll(~cm & lt_mask);
                pend_ci = (uint32_t)__builtin_amdgcn_ds_permute((int)(dst * 4), (int)cj);
                pend_n = uniform(np);
                pend_off = 0;
                continue;
            }
            const uint32_t navail = pend_n - pend_off;
            const uint32_t rank = (uint32_t)__popcll(idle & lt_mask);
            const bool take = ((idle >> lane) & 1) && rank < navail;
            uint32_t got = (uint32_t)__builtin_amdgcn_ds_bpermute((int)((pend_off + (take ? rank : 0u)) * 4), (int)pend_ci);
            got = take ? got : ~0u;
            got = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(seg_lane0 * 4), (int)got);
            pend_off = uniform(pend_off + min(navail, (uint32_t)__popcll(idle)));
            if (lane_ok && !busy && got != ~0u) {
                ci = got;
                const uint64_t o = a.coff[ci];
                n = (uint32_t)(a.coff[ci + 1] - o);
                T = a.cchars + o;
                const int32_t d = (int32_t)m - (int32_t)n;
                const int32_t s = -(int32_t)((uint32_t)((int32_t)k - d) >> 1); // dlo
                g0 = s - 1 + 64 * (int32_t)(WPL * L + PAD);
#pragma unroll
                for (int i = 0; i < WPL; ++i) {
                    const int32_t wd = (int32_t)(WPL * L) + i; // this word of the window
                    const int32_t dL = s + 64 * wd + 62, dt = dL - 63;
                    const int32_t gmin = dt <= d ? d : 2 * dt - d;
                    scmax[i] = (int32_t)k + dL - gmin;
                    sc[i] = dL > 0 ? dL : 0;
                    const int32_t sh = 1 - s - 64 * wd; // first window bit at row >= 1
                    VP[i] = sh <= 0 ? ~0ull : (sh >= 64 ? 0ull : (~0ull << sh));
                    VN[i] = 0;
                }
                const int32_t bD = d - s + 1;
                isD = (uint32_t)(bD >> 6) / WPL == L;
                iD = (uint32_t)(bD >> 6) % WPL;
                dmask = (~0ull << ((uint32_t)bD & 63u)) & (~0ull >> 1);
                jt = 0;
                busy = true;
            }
        }
        if (ballot(busy) == 0) break;

        // ---- 8 text units of every running pair ----
        uint32_t u[8], ro[8];
        load_units<uint8_t, 8>(busy ? T + jt : P, u);
#pragma unroll
        for (int q = 0; q < 8; ++q) ro[q] = rowmap[u[q]];
        // idle lanes step too (their segments' state is reset when they take a pair, and
        // the carries and the shifted bit stay inside segments); only a pair that ends
        // inside this chunk must keep its state at its last text unit
        if (ballot(busy && jt + 8 > n) == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                band_step<MULTI, false, WPL>((const uint8_t*)smem_band, ro[q], (uint32_t)(g0 + (int32_t)jt + q), LASTM,
                                             lane, last, VP, VN, sc, true);
        }
        else {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                band_step<MULTI, true, WPL>((const uint8_t*)smem_band, ro[q], (uint32_t)(g0 + (int32_t)jt + q), LASTM,
                                            lane, last, VP, VN, sc, jt + (uint32_t)q < n);
        }
        if (busy) jt += 8;
        bool lane_alive = false;
#pragma unroll
        for (int i = 0; i < WPL; ++i) lane_alive = lane_alive || sc[i] <= scmax[i];
        const uint64_t alive = ballot(busy && lane_alive);
        const bool seg_alive = ((alive >> seg_lane0) & seg_mask) != 0;
        const bool fin = busy && jt >= n;
        const bool dead = busy && !fin && !seg_alive;
        if ((fin && isD) || (dead && L == 0)) {
            // D[m][n]: the word's bottom cell less the vertical deltas of the rows below row m
            // (window bits d - dlo + 1 .. 62 of that word in the final column)
            int32_t scD = sc[0];
            uint64_t VPD = VP[0], VND = VN[0];
#pragma unroll
            for (int i = 1; i < WPL; ++i) {
                if (iD == (uint32_t)i) {
                    scD = sc[i];
                    VPD = VP[i];
                    VND = VN[i];
                }
            }
            const int32_t score = scD - __popcll(VPD & dmask) + __popcll(VND & dmask);
            const uint64_t raw = fin && score <= (int32_t)k ? (uint64_t)score : (uint64_t)k + 1;
            out[(uint64_t)qi * a.out_stride + ci] = finalize_int(spec, m, n, raw);
        }
        if ((fin || dead) && L == 0) {
            cnt_steps += (uint64_t)min(jt, n) * B;
            ++cnt_pairs;
        }
        if (fin || dead) busy = false;
    }
    if (a.counters) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            cnt_steps += __shfl_xor(cnt_steps, off, 64);
            cnt_pairs += __shfl_xor(cnt_pairs, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&a.counters[0], (unsigned long long)cnt_steps);
            atomicAdd(&a.counters[1], (unsigned long long)cnt_pairs);
        }
    }
}

// MULTI: some pair of the launch has a window of several words.  WPL2: some query runs
// two words per lane; each workgroup then takes the words per lane the host chose for its
// query (a.wpl[slot]).  (Launches without such queries keep the smaller register footprint
// of the one-word body.)
template <bool MULTI, bool WPL2>
__global__ __launch_bounds__(BAND_THREADS) void k_band(BandArgs a)
{
    extern __shared__ uint32_t smem_band[];
    if constexpr (WPL2) {
        if (a.wpl[blockIdx.x / a.chunks_per_query] == 2) {
            band_run<true, 2>(a, smem_band);
            return;
        }
    }
    band_run<MULTI, 1>(a, smem_band);
}

hipError_t launch_band(const BandArgs& a, uint32_t nblocks, uint32_t lds_bytes, bool multi, bool wpl2,
                       uint32_t threads, hipStream_t s)
{
    if (wpl2)
        hipLaunchKernelGGL((k_band<true, true>), dim3(nblocks), dim3(threads), lds_bytes, s, a);
    else if (multi)
        hipLaunchKernelGGL((k_band<true, false>), dim3(nblocks), dim3(threads), lds_bytes, s, a);
    else
        hipLaunchKernelGGL((k_band<false, false>), dim3(nblocks), dim3(threads), lds_bytes, s, a);
    return hipGetLastError();
}

} // namespace rfa
