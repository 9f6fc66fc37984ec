// rfa_wlev.hip -- Levenshtein distance with general weights: the generalized
// Wagner-Fischer recurrence (Levenshtein_impl.hpp:85-111), which
// CachedLevenshtein::_distance (Levenshtein.hpp:445-477) runs unless
// insert == delete and replace is either equal to them or at least their sum.
//
// D(i, j) is the cost of turning s1[0, i) into s2[0, j) (s1 the query, s2 the choice):
//   D(i, 0) = i * del,  D(0, j) = j * ins,
//   D(i, j) = D(i-1, j-1)                                               s1[i-1] == s2[j-1]
//           = min(D(i-1, j) + del, D(i, j-1) + ins, D(i-1, j-1) + rep)  otherwise.
// The kernels compute D(|s1|, |s2|) exactly; the reference's min_edits test and
// the cut (dist <= max ? dist : max + 1) are applied by the shared semantics.
// Two facts shape them:
//   * replace costs above ins + del are never used (a delete plus an insert is
//     cheaper), so rep is clamped to ins + del, which keeps every term small;
//   * swapping s1 and s2 together with ins and del gives the same value, which
//     lets the long kernel hold the shorter side across its lanes.
// All values stay below 2^30 (the host checks the weights against the lengths),
// so the recurrence runs in 32-bit integers.
//
// Tile kernel (queries of at most 64 units, one per lane, the choice's units
// uniform across the wave): each lane keeps its query's column D(0..m, j) in
// registers and the tile's PM row of the choice unit gives the match bits, so a
// cell costs an extract, a multiply-add and a three-way minimum.
// Long kernel (longer queries, one wave per pair): the pattern's rows split over
// the lanes, K consecutive rows each.  One column step is
//   A(i)   = min(D(i, j-1) + ins, D(i-1, j-1) + cost(i, j))
//   D(i,j) = min over k <= i of A(k) + (i - k) * del        (A(0) = D(0, j))
// i.e. a prefix minimum of A(k) - k * del: serial within a lane, a wave scan
// across the lanes.
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

namespace {

__device__ __forceinline__ uint32_t wl_wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return uniform(v);
}

// a * b + c on the 24-bit multiplier (a is a 0/1 match bit, b < 2^24).  Written as
// the instruction itself: the compiler would otherwise turn the multiply by a bit
// into a 64-bit compare and a select.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// bit i of the 64 (or 32) mismatch bits held as two 32-bit halves
template <uint32_t I>
__device__ __forceinline__ uint32_t miss_bit(uint32_t ylo, uint32_t yhi)
{
    return __builtin_amdgcn_ubfe(I < 32 ? ylo : yhi, I & 31, 1);
}

// One column step of the tile kernel: D(., j-1) -> D(., j).  Y has bit i - 1
// set where s1[i-1] differs from the choice unit; rows above mt are left alone
// (no query of the wave reaches them).  Per cell: a bit extract, a multiply-add,
// two adds and a three-way minimum.  The diagonal term of row i + 1 is taken from
// D(i, j-1) before D(i, j) overwrites it, so every row updates its own register
// (no rotation of the column, no copies where the step ends early).
template <typename W, uint32_t BITS, uint32_t I = 1>
__device__ __forceinline__ void wf_rows(uint32_t (&D)[BITS + 1], uint32_t ylo, uint32_t yhi, uint32_t ins,
                                        uint32_t del, uint32_t rep, uint32_t mt, uint32_t sub)
{
    if constexpr (I <= BITS) {
        if ((I & 3) == 1 && I > mt) return;
        const uint32_t up = D[I];
        uint32_t nsub = 0;
        if constexpr (I < BITS) nsub = mad_u24(miss_bit<I>(ylo, yhi), rep, up);
        D[I] = min(min(up + ins, sub), D[I - 1] + del);
        wf_rows<W, BITS, I + 1>(D, ylo, yhi, ins, del, rep, mt, nsub);
    }
}

template <typename W, uint32_t BITS>
__device__ __forceinline__ void wf_step(uint32_t (&D)[BITS + 1], W Y, uint32_t ins, uint32_t del, uint32_t rep,
                                        uint32_t mt)
{
    const uint32_t ylo = (uint32_t)Y;
    const uint32_t yhi = BITS > 32 ? (uint32_t)((uint64_t)Y >> 32) : 0u;
    const uint32_t sub = mad_u24(miss_bit<0>(ylo, yhi), rep, D[0]);
    D[0] += ins;
    wf_rows<W, BITS>(D, ylo, yhi, ins, del, rep, mt, sub);
}

} // namespace

// ---- tile kernel --------------------------------------------------------------

template <typename CharT, typename W, bool INTK>
__global__ __launch_bounds__(TILE_THREADS) void k_wtile(TileArgs a)
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
    ScoreSpec spec = a.spec;
    if (a.qcut && qact) spec.cutoff_u = a.qcut[qidx];
    const uint32_t ins = (uint32_t)spec.ins, del = (uint32_t)spec.del;
    const uint32_t rep = (uint32_t)u64min(spec.rep, spec.ins + spec.del);
    const uint32_t mt = wl_wave_max(m);
    const W* __restrict__ col = pm + lane;

    const uint32_t it0 = blockIdx.x * a.choices_per_block;
    const uint32_t it1 = min(a.ncsel, it0 + a.choices_per_block);
    constexpr uint32_t G = 4;
    for (uint32_t g0 = it0 + uniform(wave) * G; g0 < it1; g0 += TILE_WAVES * G) {
        const uint32_t gn = min(G, it1 - g0);
        uint64_t outv[G];
        for (uint32_t t = 0; t < gn; ++t) {
            const uint32_t cidx = a.corder ? a.corder[g0 + t] : g0 + t;
            const uint64_t co0 = a.coff[cidx];
            const uint32_t n = uniform((uint32_t)(a.coff[cidx + 1] - co0));
            const uint64_t co = ((uint64_t)uniform((uint32_t)(co0 >> 32)) << 32) | uniform((uint32_t)co0);
            uint32_t D[BITS + 1];
#pragma unroll
            for (uint32_t i = 0; i <= BITS; ++i) D[i] = i * del;
            uint32_t unit_next = lane < n ? (uint32_t)c[co + lane] : 0u;
            for (uint32_t k0 = 0; k0 < n; k0 += 64) {
                const uint32_t row = k0 + lane < n ? look(unit_next) : 0u;
                if (k0 + 64 + lane < n) unit_next = (uint32_t)c[co + k0 + 64 + lane];
                const uint32_t steps = min(64u, n - k0);
                W Xn = col[readlane(row, 0) * 64];
                for (uint32_t j = 0; j < steps; ++j) {
                    const W X = Xn;
                    Xn = col[readlane(row, min(j + 1, 63u)) * 64];
                    wf_step<W, BITS>(D, (W)~X, ins, del, rep, mt);
                }
            }
            uint32_t res = D[0];
#pragma unroll
            for (uint32_t i = 1; i <= BITS; ++i) res = (i == m) ? D[i] : res;
            RawResult r{};
            r.v = res;
            if constexpr (INTK)
                outv[t] = finalize_int(spec, m, n, res);
            else
                outv[t] = finalize_bits(spec, m, n, r);
        }
        if (qact) {
            uint64_t* dst = (uint64_t*)a.out + (uint64_t)qidx * a.out_stride;
            if (a.corder) {
                for (uint32_t t = 0; t < gn; ++t) dst[a.corder[g0 + t]] = outv[t];
            }
            else if (gn == G && ((uintptr_t)(dst + g0) & 15) == 0) {
                uint4* d4 = (uint4*)(dst + g0);
#pragma unroll
                for (uint32_t t = 0; t < G; t += 2)
                    d4[t / 2] = make_uint4((uint32_t)outv[t], (uint32_t)(outv[t] >> 32), (uint32_t)outv[t + 1],
                                           (uint32_t)(outv[t + 1] >> 32));
            }
            else {
                for (uint32_t t = 0; t < gn; ++t) dst[g0 + t] = outv[t];
            }
        }
    }
}

// ---- long kernel: one wave per pair, the pattern's rows across the lanes ------
//
// The pattern's rows sit K per lane (64 K rows); a column step takes the text unit
// wave-uniform, forms each cell's candidates and closes the column's vertical
// (insertion along the pattern) chain with a min-plus scan across the lanes.  A
// pattern longer than 64 K rows (both strings over 1,024 units) runs in strips of
// 64 K rows: each strip reads its top boundary, the previous strip's bottom row,
// from the wave's boundary row in scratch and writes its own bottom row there, 64
// columns per coalesced load and store.

template <typename CharT, int K>
__global__ __launch_bounds__(256) void k_wlong(WlevLongArgs a)
{
    const CharT* __restrict__ qs = (const CharT*)a.qchars;
    const CharT* __restrict__ cs = (const CharT*)a.cchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    const uint64_t items = (uint64_t)a.nql * a.nc;
    constexpr int32_t INF = 0x7fffffff;
    constexpr uint32_t SROWS = 64 * K; // pattern rows per strip
    int32_t* __restrict__ bnd = a.bnd ? a.bnd + wave0 * a.bnd_stride : nullptr;

    for (uint64_t it = wave0; it < items; it += nwaves) {
        const uint32_t qi = a.qlist[it / a.nc];
        const uint32_t ci = (uint32_t)(it % a.nc);
        const uint64_t qo = a.qoff[qi], co = a.coff[ci];
        const uint32_t m = uniform((uint32_t)(a.qoff[qi + 1] - qo));
        const uint32_t n = uniform((uint32_t)(a.coff[ci + 1] - co));
        ScoreSpec spec = a.spec;
        if (a.qcut) spec.cutoff_u = a.qcut[qi];
        // the pattern (rows) is the query unless it is longer than a strip and the choice is not
        const bool swap = m > SROWS && n <= SROWS;
        const CharT* P = swap ? cs + co : qs + qo;
        const CharT* T = swap ? qs + qo : cs + co;
        const uint32_t lp = swap ? n : m, lt = swap ? m : n;
        const int32_t rc = (int32_t)(swap ? spec.ins : spec.del); // cost of a pattern row
        const int32_t cc = (int32_t)(swap ? spec.del : spec.ins); // cost of a text column
        const int32_t rep = (int32_t)u64min(spec.rep, spec.ins + spec.del);
        const uint32_t nstrip = lp > SROWS ? (lp + SROWS - 1) / SROWS : 1u; // > 1 only with a boundary row

        int32_t V[K];
        for (uint32_t strip = 0; strip < nstrip; ++strip) {
            const uint32_t row0 = strip * SROWS; // rows above the strip
            const bool last = strip + 1 == nstrip;
            uint32_t pu[K];
            int32_t rr[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t r = row0 + lane * K + k + 1; // row of cell k
                pu[k] = r <= lp ? (uint32_t)P[r - 1] : 0u;
                rr[k] = (int32_t)r * rc;
                V[k] = rr[k];
            }
            int32_t top = (int32_t)row0 * rc; // D[row0][j], j = 0 first
            for (uint32_t j0 = 0; j0 < lt; j0 += 64) {
                const uint32_t tu = j0 + lane < lt ? (uint32_t)T[j0 + lane] : 0u;
                const uint32_t steps = min(64u, lt - j0);
                // the top boundary at columns j0 + 1 .. j0 + 64
                int32_t tb = (int32_t)(j0 + lane + 1) * cc;
                if (strip > 0) tb = j0 + lane < lt ? bnd[j0 + lane + 1] : 0;
                int32_t nb = 0; // this strip's bottom row at column j0 + lane + 1
                for (uint32_t s = 0; s < steps; ++s) {
                    const uint32_t t = readlane(tu, s);
                    const int32_t top_old = top, top_new = (int32_t)readlane((uint32_t)tb, s);
                    top = top_new;
                    int32_t prev = __shfl_up(V[K - 1], 1, 64);
                    if (lane == 0) prev = top_old;
                    int32_t run = INF;
                    int32_t Pk[K];
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const int32_t A = min(V[k] + cc, prev + (pu[k] == t ? 0 : rep));
                        prev = V[k];
                        run = min(run, A - rr[k]);
                        Pk[k] = run;
                    }
                    int32_t incl = run;
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const int32_t o = __shfl_up(incl, off, 64);
                        if ((int)lane >= off) incl = min(incl, o);
                    }
                    int32_t carry = __shfl_up(incl, 1, 64);
                    if (lane == 0) carry = INF;
                    // the scan runs on values less their row's rr: the top boundary less row0's
                    carry = min(carry, top_new - (int32_t)row0 * rc);
#pragma unroll
                    for (int k = 0; k < K; ++k) V[k] = min(carry, Pk[k]) + rr[k];
                    if (!last) { // a full strip: its bottom row is lane 63's last cell
                        const int32_t bv = (int32_t)readlane((uint32_t)V[K - 1], 63);
                        nb = lane == s ? bv : nb;
                    }
                }
                if (!last && lane < steps) bnd[j0 + lane + 1] = nb;
            }
        }
        uint32_t res;
        if (lp == 0) {
            res = lt * (uint32_t)cc;
        }
        else {
            const uint32_t lr = (lp - 1) % SROWS; // the last row's place in the last strip
            const uint32_t kk = lr % K;
            int32_t v = V[0];
#pragma unroll
            for (int k = 1; k < K; ++k) v = ((uint32_t)k == kk) ? V[k] : v;
            res = (uint32_t)readlane((uint32_t)v, lr / K);
        }
        RawResult r{};
        r.v = res;
        const uint64_t bits = finalize_bits(spec, m, n, r);
        if (lane == 0) ((uint64_t*)a.out)[(uint64_t)qi * a.out_stride + ci] = bits;
    }
}

// ---- launches -------------------------------------------------------------------

template <typename CharT, typename W>
static void launch_wtile_w(const TileArgs& a, uint32_t ntiles, uint32_t nbx, size_t lds, hipStream_t s)
{
    const bool intk = a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY;
    if (intk)
        hipLaunchKernelGGL((k_wtile<CharT, W, true>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a);
    else
        hipLaunchKernelGGL((k_wtile<CharT, W, false>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a);
}

hipError_t launch_wtile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                        hipStream_t s)
{
    const size_t lds = tile_lds_bytes(char_width, word_bits, a.rows_cap, a.hash_bits);
    if (word_bits == 32) {
        if (char_width == 1)
            launch_wtile_w<uint8_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_wtile_w<uint16_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_wtile_w<uint32_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
    }
    else {
        if (char_width == 1)
            launch_wtile_w<uint8_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_wtile_w<uint16_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_wtile_w<uint32_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
    }
    return hipGetLastError();
}

template <typename CharT>
static void launch_wlong_c(uint32_t cells, const WlevLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (cells <= 1)
        hipLaunchKernelGGL((k_wlong<CharT, 1>), dim3(nblocks), dim3(256), 0, s, a);
    else if (cells <= 2)
        hipLaunchKernelGGL((k_wlong<CharT, 2>), dim3(nblocks), dim3(256), 0, s, a);
    else if (cells <= 4)
        hipLaunchKernelGGL((k_wlong<CharT, 4>), dim3(nblocks), dim3(256), 0, s, a);
    else if (cells <= 8)
        hipLaunchKernelGGL((k_wlong<CharT, 8>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_wlong<CharT, 16>), dim3(nblocks), dim3(256), 0, s, a);
}

hipError_t launch_wlong(int char_width, uint32_t cells_per_lane, const WlevLongArgs& a, uint32_t nblocks,
                        hipStream_t s)
{
    if (cells_per_lane > WLEV_LONG_MAX / 64) return hipErrorInvalidValue;
    if (char_width == 1)
        launch_wlong_c<uint8_t>(cells_per_lane, a, nblocks, s);
    else if (char_width == 2)
        launch_wlong_c<uint16_t>(cells_per_lane, a, nblocks, s);
    else
        launch_wlong_c<uint32_t>(cells_per_lane, a, nblocks, s);
    return hipGetLastError();
}

} // namespace rfa
