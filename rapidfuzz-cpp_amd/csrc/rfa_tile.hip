// rfa_tile.hip -- query-tile kernel: 64 short queries (one per lane) against a
// stream of choices that is uniform across the wave.
//
// Shape.  A workgroup of 8 waves owns one tile of up to 64 queries, each at
// most 64 code units (the reference's single-word paths:
// Levenshtein_impl.hpp:259-322, LCSseq_impl.hpp:205-260 with N=1,
// OSA_impl.hpp:34-74, Jaro_impl.hpp:94-132 and 258-276).  The tile's
// pattern-match vectors PM[symbol][lane] are built into LDS once
// (PatternMatchVector.hpp:62-134 semantics; symbols compacted to dense rows),
// then each wave walks its share of the choices.  The current choice's code
// units are the same for every lane, so PM[row][lane] is one contiguous,
// conflict-free 64-lane LDS read per text position and the whole wave runs
// the same trip count.
//
// Word width W is 32 bits for tiles whose queries all fit in 32 units and 64
// bits otherwise (a 64-bit op is two VALU ops on CDNA).
#include <type_traits>

#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

// One pass of a single-word recurrence over cnt (1..64) text positions whose PM
// rows are `row` of lanes 0..cnt-1; step(X, i) runs position i.  Whole groups
// of 4 positions run with the loads two positions ahead of the recurrence (the
// scheduler kept from pulling a load down to its use), so an LDS read's latency
// overlaps two steps of the wave's own work; the last 0-3 positions follow.
// Loads past cnt read row 63's word and are never stepped.
// stop(done) is asked after each whole group with the number of positions
// done so far; true ends the pass (the rest of the positions are not stepped)
// and the function returns true.
template <typename W, typename Step, typename Stop>
__device__ __forceinline__ bool pipelined_rows(const W* __restrict__ col, uint32_t row, uint32_t cnt, Step& step,
                                               Stop& stop)
{
    auto ld = [&](uint32_t i) -> W { return col[readlane(row, i < 63u ? i : 63u) * 64]; };
    W x0 = ld(0), x1 = ld(1);
    const uint32_t full = cnt & ~3u;
    uint32_t i = 0;
    for (; i < full; i += 4) {
        W x2 = ld(i + 2), x3 = ld(i + 3);
        __builtin_amdgcn_sched_barrier(0);
        step(x0, i);
        step(x1, i + 1);
        x0 = ld(i + 4);
        x1 = ld(i + 5);
        __builtin_amdgcn_sched_barrier(0);
        step(x2, i + 2);
        step(x3, i + 3);
        if (stop(i + 4)) return true;
    }
    if (i < cnt) {
        step(x0, i);
        if (i + 1 < cnt) {
            step(x1, i + 1);
            if (i + 2 < cnt) step(ld(i + 2), i + 2);
        }
    }
    return false;
}

template <typename W, typename Step>
__device__ __forceinline__ void pipelined_rows(const W* __restrict__ col, uint32_t row, uint32_t cnt, Step& step)
{
    auto never = [](uint32_t) { return false; };
    pipelined_rows<W>(col, row, cnt, step, never);
}

constexpr uint32_t T_AND_ANDN = (TA & TB & ~TC) & 0xFF; // a & b & ~c
constexpr uint32_t T_AND_OR = ((TA & TB) | TC) & 0xFF;  // (a & b) | c
constexpr uint32_t T_AND3 = (TA & TB & TC) & 0xFF;      // a & b & c
constexpr uint32_t T_ANDN2 = (TA & ~TB) & 0xFF;         // a & ~b

// Jaro word path (one query of <= 64 units per lane, a choice of n <= 64 units
// uniform across the wave): flag_similar_characters_word and
// count_transpositions_word (Jaro_impl.hpp:94-132, 258-276), plus the common
// prefix (at most 4) that Jaro-Winkler boosts with (JaroWinkler_impl.hpp:41-50).
//
// Flagging: text position j takes the lowest unflagged query position of its
// symbol within [j - Bound, j + Bound].  BoundMask starts as the Bound + 1 low
// bits and moves up one bit per position, growing while j < Bound.
// Transpositions: the flagged text positions in order are paired with the
// flagged query positions in order; a pair whose text symbol does not occur at
// the query position counts.  64-bit words run as 32-bit halves; a 64-bit
// shift-left-with-carry-in is two add-with-carry instructions.
template <typename W>
__device__ __forceinline__ RawResult jaro_word(const W* __restrict__ col, uint32_t row, uint32_t m, uint32_t n)
{
    RawResult res{0, 0, 0};
    uint32_t Bound = max(m, n) / 2;
    if (Bound > 0) Bound--;
    constexpr uint32_t BITS = sizeof(W) * 8;
    W BM = (Bound + 1 >= BITS) ? ~(W)0 : (((W)1 << (Bound + 1)) - 1);
    uint32_t Pl = 0, Ph = 0; // flagged query positions
    uint64_t T = 0;          // flagged text positions, position j at bit n - 1 - j
    uint32_t pfx = 0;
    const uint32_t np = uniform(min(n, 4u));
    for (uint32_t j = 0; j < np; ++j) {
        const W PMr = col[readlane(row, j) * 64];
        if (pfx == j && j < m && ((PMr >> j) & 1)) pfx = j + 1;
    }
    auto flag = [&](W X, uint32_t j) {
        const W grow = j < Bound;
        uint32_t hit;
        if constexpr (sizeof(W) == 4) {
            const uint32_t x = bop3<T_AND_ANDN>((uint32_t)X, (uint32_t)BM, Pl);
            Pl = bop3<T_AND_OR>(x, 0u - x, Pl);
            hit = x != 0;
        }
        else {
            const uint32_t xl = bop3<T_AND_ANDN>((uint32_t)X, (uint32_t)BM, Pl);
            const uint32_t xh = bop3<T_AND_ANDN>((uint32_t)((uint64_t)X >> 32), (uint32_t)((uint64_t)BM >> 32), Ph);
            const uint64_t nx = 0ull - (((uint64_t)xh << 32) | xl);
            Pl = bop3<T_AND_OR>(xl, (uint32_t)nx, Pl);
            Ph = bop3<T_AND_OR>(xh, (uint32_t)(nx >> 32), Ph);
            hit = (xl | xh) != 0;
        }
        T = (T << 1) + hit;
        BM = (W)((BM << 1) + grow);
    };
    if (n) pipelined_rows<W>(col, row, n, flag);

    uint32_t tr = 0;
    uint32_t Ql = Pl, Qh = Ph; // flagged query positions not yet paired
    const uint32_t Tl = (uint32_t)T, Th = (uint32_t)(T >> 32);
    auto walk = [&](W X, uint32_t j) {
        // all ones when text position j is flagged, else zero
        const uint32_t b = n - 1 - j;
        const uint32_t tm = (uint32_t)((int32_t)((b < 32 ? Tl : Th) << (31 - (b & 31))) >> 31);
        if constexpr (sizeof(W) == 4) {
            const uint32_t sel = bop3<T_AND3>(Ql, 0u - Ql, tm); // the lowest unpaired query position
            tr += bop3<T_ANDN2>(sel, (uint32_t)X, 0u) != 0;
            Ql ^= sel;
        }
        else {
            const uint64_t nq = 0ull - (((uint64_t)Qh << 32) | Ql);
            const uint32_t sl = bop3<T_AND3>(Ql, (uint32_t)nq, tm);
            const uint32_t sh = bop3<T_AND3>(Qh, (uint32_t)(nq >> 32), tm);
            tr += (bop3<T_ANDN2>(sl, (uint32_t)X, 0u) | bop3<T_ANDN2>(sh, (uint32_t)((uint64_t)X >> 32), 0u)) != 0;
            Ql ^= sl;
            Qh ^= sh;
        }
    };
    // the walk is needed only when some lane flagged a text position
    if (ballot(T != 0)) pipelined_rows<W>(col, row, n, walk);
    res.v = (uint64_t)(__popc(Pl) + __popc(Ph));
    res.tr = tr;
    res.prefix = pfx;
    return res;
}

// EPI selects the epilogue compiled in: EPI_INT for the integer-valued kinds
// (which keeps the register footprint small), EPI_NORM for the normalized
// kinds whose launch-level choices NormEpi makes once (Cached Levenshtein with
// unit weights, OSA, Indel, LCSseq, ratio), EPI_ANY for the rest.
enum : int { EPI_INT = 0, EPI_NORM = 1, EPI_ANY = 2 };

template <int ALGO, int EPI>
__device__ __forceinline__ uint64_t tile_result(const ScoreSpec& spec, const NormEpi& ne, uint32_t m, uint32_t n,
                                                RawResult r)
{
    if constexpr (EPI == EPI_INT)
        return finalize_int(spec, m, n, r.v);
    else if constexpr (EPI == EPI_NORM) {
        const uint32_t M = norm_epi_max(ne, m, n);
        return dbits(norm_epi_value(ne, M, norm_epi_dist(ne, M, (uint32_t)r.v), norm_epi_cut(ne, M)));
    }
    else
        return finalize_bits_algo<ALGO>(spec, m, n, r);
}

template <int ALGO, int EPI>
__device__ __forceinline__ bool tile_settled(const ScoreSpec& spec, const NormEpi& ne, uint32_t m, uint32_t n)
{
    const uint64_t lo = raw_range_lo<ALGO>(m, n), hi = raw_range_hi<ALGO>(m, n);
    return tile_result<ALGO, EPI>(spec, ne, m, n, RawResult{lo, 0, 0}) ==
           tile_result<ALGO, EPI>(spec, ne, m, n, RawResult{hi, 0, 0});
}

template <int ALGO, typename CharT, typename W, int EPI>
__global__ __launch_bounds__(TILE_THREADS) void k_tile(TileArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    W* pm = (W*)smem_u64;
    uint8_t* extra = (uint8_t*)(pm + (size_t)a.rows_cap * 64);
    const CharT* __restrict__ q = (const CharT*)a.qchars;
    const CharT* __restrict__ c = (const CharT*)a.cchars;

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
    NormEpi ne{};
    if constexpr (EPI == EPI_NORM) norm_epi_setup(a.spec, ne); // the host checked that it applies

    // Choice lengths whose result this lane's length alone decides (bit n of
    // settle_lo for n < 64, settle_64 for n = 64; longer choices are tested
    // when they come).  A choice that every lane settles skips the recurrence.
    // The waves share the work: wave w tests n = w, w + 8, ..., and the bits
    // meet in LDS.
    uint64_t settle_lo = 0;
    bool settle_64 = false;
    if constexpr (ALGO != ALGO_JARO) {
        if (a.settle) {
            uint64_t* sl = (uint64_t*)(extra + tile_extra_bytes(sizeof(CharT), a.hash_bits));
            if (wave == 0) sl[lane] = sl[64 + lane] = 0;
            __syncthreads();
            uint64_t mine = 0;
#pragma unroll 1
            for (uint32_t n = wave; n < 64; n += TILE_WAVES)
                mine |= (uint64_t)tile_settled<ALGO, EPI>(spec, ne, m, n) << n;
            atomicOr((unsigned long long*)&sl[lane], (unsigned long long)mine);
            if (wave == TILE_WAVES - 1) sl[64 + lane] = tile_settled<ALGO, EPI>(spec, ne, m, 64);
            __syncthreads();
            settle_lo = sl[lane];
            settle_64 = sl[64 + lane] != 0;
        }
    }

    constexpr uint32_t BITS = sizeof(W) * 8;
    const uint32_t it0 = blockIdx.x * a.choices_per_block;
    const uint32_t it1 = min(a.ncsel, it0 + a.choices_per_block);
    const W* __restrict__ col = pm + lane;

    // Each wave takes runs of G consecutive choices.  With the full choice set,
    // lane l keeps its G results for query row l in registers and stores them
    // as one contiguous 8*G-byte row segment: the output goes to HBM in whole
    // lines instead of one 8-byte word per row per choice.  A choice subset
    // (corder) writes each result as it is produced.
    //
    // Latency is taken off the critical path one choice ahead: lane t of a run
    // holds choice t's index, offset and length (loaded a run ahead), choice
    // t + 1's first 64 code units are loaded while choice t runs, and its PM
    // rows are looked up before choice t's epilogue.
    // Integer results are held as 32 bits (the host routes launches whose values could
    // reach 2^32 to EPI_ANY), 16 per lane: one 128-byte row segment per run.  Doubles
    // are held 16 per lane in the normalized variant (128-byte segments again), 8 in the
    // general one.
    constexpr uint32_t G = EPI == EPI_ANY ? 8 : 16;
    using OutT = std::conditional_t<EPI == EPI_INT, uint32_t, uint64_t>;
    const bool direct = a.corder != nullptr;
    uint32_t g0 = it0 + uniform(wave) * G;
    if (g0 >= it1) return;

    auto run_load = [&](uint32_t r0, uint32_t& ci, uint32_t& olo, uint32_t& ohi, uint32_t& len) {
        ci = 0, olo = 0, ohi = 0, len = 0;
        const uint32_t it = r0 + lane;
        if (lane < G && it < it1) {
            ci = direct ? a.corder[it] : it;
            const uint64_t o = a.coff[ci];
            olo = (uint32_t)o, ohi = (uint32_t)(o >> 32);
            len = (uint32_t)(a.coff[ci + 1] - o);
        }
    };
    auto text_unit = [&](uint64_t o, uint32_t len) -> uint32_t {
        uint32_t u = 0;
        if (lane < len) u = (uint32_t)c[o + lane];
        return u;
    };
    auto text_row = [&](uint32_t u, uint32_t len) -> uint32_t { return lane < len ? look(u) : 0u; };

    uint32_t r_ci, r_olo, r_ohi, r_len; // the current run
    run_load(g0, r_ci, r_olo, r_ohi, r_len);
    uint32_t gn = min(G, it1 - g0);
    uint32_t n_cur = uniform(readlane(r_len, 0));
    uint64_t co_cur = ((uint64_t)readlane(r_ohi, 0) << 32) | readlane(r_olo, 0);
    uint32_t row_cur = text_row(text_unit(co_cur, n_cur), n_cur);

    for (;;) {
        const uint32_t g1 = g0 + TILE_WAVES * G; // this wave's next run
        uint32_t x_ci, x_olo, x_ohi, x_len;
        run_load(g1, x_ci, x_olo, x_ohi, x_len);
        OutT outv[G];
        for (uint32_t t = 0; t < gn; ++t) {
            const uint32_t cidx = uniform(readlane(r_ci, t));
            const uint32_t n = n_cur;
            const uint64_t co = co_cur;
            const uint32_t row = row_cur;
            // the next choice of this wave: its length, offset and first units
            uint32_t n_nx = 0;
            uint64_t co_nx = 0;
            if (t + 1 < gn) {
                n_nx = uniform(readlane(r_len, t + 1));
                co_nx = ((uint64_t)readlane(r_ohi, t + 1) << 32) | readlane(r_olo, t + 1);
            }
            else if (g1 < it1) {
                n_nx = uniform(readlane(x_len, 0));
                co_nx = ((uint64_t)readlane(x_ohi, 0) << 32) | readlane(x_olo, 0);
            }
            const uint32_t u_nx = text_unit(co_nx, n_nx);

            RawResult res{0, 0, 0};
            bool skip = false;
            if constexpr (ALGO != ALGO_JARO) {
                if (a.settle) {
                    const bool st = n < 64 ? ((settle_lo >> n) & 1) != 0
                                           : (n == 64 ? settle_64 : tile_settled<ALGO, EPI>(spec, ne, m, n));
                    skip = ballot(!st) == 0;
                    if (skip) res.v = raw_range_lo<ALGO>(m, n); // any value of the range gives this result
                }
            }

            if (skip) {
            }
            else if constexpr (ALGO == ALGO_LEV || ALGO == ALGO_OSA || ALGO == ALGO_LCS) {
                // state as 32-bit halves (h* unused for 32-bit words)
                uint32_t VPl = ~0u, VPh = ~0u, VNl = 0, VNh = 0, D0l = 0, D0h = 0, PMl = 0, PMh = 0;
                auto step = [&](W X, uint32_t) {
                    if constexpr (sizeof(W) == 4) {
                        if constexpr (ALGO == ALGO_LEV)
                            lev_step((uint32_t)X, VPl, VNl);
                        else if constexpr (ALGO == ALGO_OSA)
                            osa_step((uint32_t)X, VPl, VNl, D0l, PMl);
                        else
                            lcs_step((uint32_t)X, VPl);
                    }
                    else {
                        if constexpr (ALGO == ALGO_LEV)
                            lev_step((uint64_t)X, VPl, VPh, VNl, VNh);
                        else if constexpr (ALGO == ALGO_OSA)
                            osa_step((uint64_t)X, VPl, VPh, VNl, VNh, D0l, D0h, PMl, PMh);
                        else
                            lcs_step((uint64_t)X, VPl, VPh);
                    }
                };
                // LCS: when no lane can reach its threshold any more (the LCS so far
                // plus the positions left), every result is already the cut value
                uint32_t need = 0;
                bool can_stop = false;
                if constexpr (ALGO == ALGO_LCS) {
                    if (a.settle) {
                        if constexpr (EPI == EPI_NORM) {
                            // an LCS below this gives a distance above the cut
                            const uint32_t M = norm_epi_max(ne, m, n), cd = norm_epi_cut(ne, M);
                            need = M <= cd ? 0u : (ne.sum_max ? (M - cd + 1) / 2 : M - cd);
                        }
                        else {
                            need = lcs_cut_below(spec, m, n);
                        }
                        if (!qact) need = ~0u; // idle lanes never block a stop
                        can_stop = ballot(need == 0) == 0;
                    }
                }
                uint32_t k0 = 0;
                auto stop = [&](uint32_t done) -> bool {
                    if constexpr (ALGO == ALGO_LCS) {
                        if (!can_stop) return false;
                        const W S = (sizeof(W) == 8) ? (W)(((uint64_t)VPh << 32) | VPl) : (W)VPl;
                        const uint32_t bound = (uint32_t)popc<W>(~S) + (n - k0 - done);
                        return ballot(bound >= need) == 0;
                    }
                    else {
                        return false;
                    }
                };
                // (the test sits in its own copy of the loop, so runs without it pay nothing)
                bool stopped = false;
                if (can_stop) {
                    if (n) stopped = pipelined_rows<W>(col, row, min(64u, n), step, stop);
                }
                else if (n) {
                    pipelined_rows<W>(col, row, min(64u, n), step);
                }
                for (k0 = 64; k0 < n && !stopped; k0 += 64) {
                    const uint32_t cnt = uniform(min(64u, n - k0));
                    const uint32_t rk = text_row(text_unit(co + k0, cnt), cnt);
                    if (can_stop)
                        stopped = pipelined_rows<W>(col, rk, cnt, step, stop);
                    else
                        pipelined_rows<W>(col, rk, cnt, step);
                }
                const W VP = (sizeof(W) == 8) ? (W)(((uint64_t)VPh << 32) | VPl) : (W)VPl;
                const W VN = (sizeof(W) == 8) ? (W)(((uint64_t)VNh << 32) | VNl) : (W)VNl;
                if constexpr (ALGO == ALGO_LCS) {
                    res.v = (uint64_t)popc<W>(~VP); // S is kept in the VP registers
                }
                else {
                    // D[m][n] = D[0][n] + sum of the final column's vertical deltas
                    const W M = (m >= BITS) ? ~(W)0 : (((W)1 << m) - 1);
                    res.v = m ? (uint64_t)n + popc<W>(VP & M) - popc<W>(VN & M) : n;
                }
            }
            else {
                res = jaro_word<W>(col, row, m, n);
            }
            const uint32_t row_nx = text_row(u_nx, n_nx);
            uint64_t bits;
            if constexpr (EPI == EPI_ANY) {
                if (a.pcut) {
                    ScoreSpec sp = spec;
                    if (qact) sp.cutoff_d = a.pcut[(uint64_t)qidx * a.out_stride + cidx];
                    bits = tile_result<ALGO, EPI>(sp, ne, m, n, res);
                }
                else {
                    bits = tile_result<ALGO, EPI>(spec, ne, m, n, res);
                }
            }
            else {
                bits = tile_result<ALGO, EPI>(spec, ne, m, n, res);
            }
            if (direct) {
                if (qact) ((uint64_t*)a.out)[(uint64_t)qidx * a.out_stride + cidx] = bits;
            }
            else {
                outv[t] = (OutT)bits;
            }
            n_cur = n_nx, co_cur = co_nx, row_cur = row_nx;
        }
        if (!direct && qact) {
            uint64_t* dst = (uint64_t*)a.out + (uint64_t)qidx * a.out_stride + g0;
            if (gn == G && ((uintptr_t)dst & 15) == 0) {
                uint4* d4 = (uint4*)dst;
#pragma unroll
                for (uint32_t t = 0; t < G; t += 2)
                    d4[t / 2] = make_uint4((uint32_t)outv[t], (uint32_t)((uint64_t)outv[t] >> 32),
                                           (uint32_t)outv[t + 1], (uint32_t)((uint64_t)outv[t + 1] >> 32));
            }
            else if (gn == G) {
#pragma unroll
                for (uint32_t t = 0; t < G; ++t) dst[t] = (uint64_t)outv[t];
            }
            else {
                for (uint32_t t = 0; t < gn; ++t) dst[t] = (uint64_t)outv[t];
            }
        }
        if (g1 >= it1) break;
        g0 = g1;
        gn = min(G, it1 - g0);
        r_ci = x_ci, r_olo = x_olo, r_ohi = x_ohi, r_len = x_len;
    }
}

template <int ALGO, typename CharT, typename W>
static void launch_one(const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x, size_t lds, hipStream_t s)
{
    dim3 grid(nblocks_x, ntiles);
    const bool intk = (a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY) && a.spec.metric <= RFA_OSA;
    if constexpr (ALGO != ALGO_JARO) {
        if (intk && a.narrow_int) {
            hipLaunchKernelGGL((k_tile<ALGO, CharT, W, EPI_INT>), grid, dim3(TILE_THREADS), lds, s, a);
            return;
        }
        if (a.norm_epi) {
            hipLaunchKernelGGL((k_tile<ALGO, CharT, W, EPI_NORM>), grid, dim3(TILE_THREADS), lds, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_tile<ALGO, CharT, W, EPI_ANY>), grid, dim3(TILE_THREADS), lds, s, a);
}

template <typename CharT, typename W>
static void launch_algo(int algo, const TileArgs& a, uint32_t ntiles, uint32_t nbx, size_t lds, hipStream_t s)
{
    switch (algo) {
    case ALGO_LEV: launch_one<ALGO_LEV, CharT, W>(a, ntiles, nbx, lds, s); break;
    case ALGO_LCS: launch_one<ALGO_LCS, CharT, W>(a, ntiles, nbx, lds, s); break;
    case ALGO_OSA: launch_one<ALGO_OSA, CharT, W>(a, ntiles, nbx, lds, s); break;
    default: launch_one<ALGO_JARO, CharT, W>(a, ntiles, nbx, lds, s); break;
    }
}

size_t tile_lds_bytes(int char_width, int word_bits, uint32_t rows_cap, uint32_t hash_bits)
{
    size_t pm = (size_t)rows_cap * 64 * (word_bits / 8);
    return pm + tile_extra_bytes(char_width, hash_bits) + TILE_SETTLE_BYTES;
}

hipError_t launch_tile(int algo, int char_width, int word_bits, const TileArgs& a, uint32_t ntiles,
                       uint32_t nblocks_x, hipStream_t s)
{
    size_t lds = tile_lds_bytes(char_width, word_bits, a.rows_cap, a.hash_bits);
    if (word_bits == 32) {
        if (char_width == 1)
            launch_algo<uint8_t, uint32_t>(algo, a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_algo<uint16_t, uint32_t>(algo, a, ntiles, nblocks_x, lds, s);
        else
            launch_algo<uint32_t, uint32_t>(algo, a, ntiles, nblocks_x, lds, s);
    }
    else {
        if (char_width == 1)
            launch_algo<uint8_t, uint64_t>(algo, a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_algo<uint16_t, uint64_t>(algo, a, ntiles, nblocks_x, lds, s);
        else
            launch_algo<uint32_t, uint64_t>(algo, a, ntiles, nblocks_x, lds, s);
    }
    return hipGetLastError();
}

} // namespace rfa
