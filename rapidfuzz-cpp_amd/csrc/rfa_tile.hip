// rfa_tile.hip -- query-tile kernel: 64 short queries (one per lane) against a
// stream of choices that is uniform across the wave.
//
// Shape.  A workgroup of 4 waves owns one tile of up to 64 queries, each at
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
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

// INTK: the result kind is integer-valued (the integer epilogue alone is
// compiled in, which keeps the kernel's register footprint small).
template <int ALGO, typename CharT, typename W, bool INTK>
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

    constexpr uint32_t BITS = sizeof(W) * 8;
    const uint32_t it0 = blockIdx.x * a.choices_per_block;
    const uint32_t it1 = min(a.ncsel, it0 + a.choices_per_block);

    // Each wave takes runs of G consecutive choices.  With the full choice set,
    // lane l keeps its G results for query row l in registers and stores them
    // as one contiguous 8*G-byte row segment: the output goes to HBM in whole
    // lines instead of one 8-byte word per row per choice.  A choice subset
    // (corder) writes each result as it is produced.
    constexpr uint32_t G = INTK ? 16 : 4;
    const bool direct = a.corder != nullptr;
    for (uint32_t g0 = it0 + uniform(wave) * G; g0 < it1; g0 += TILE_WAVES * G) {
        const uint32_t gn = min(G, it1 - g0);
        uint64_t outv[G];
        for (uint32_t t = 0; t < gn; ++t) {
            const uint32_t it = g0 + t;
            const uint32_t cidx = uniform(direct ? a.corder[it] : it);
            const uint64_t co0 = a.coff[cidx];
            const uint32_t n = uniform((uint32_t)(a.coff[cidx + 1] - co0));
            const uint64_t co = ((uint64_t)uniform((uint32_t)(co0 >> 32)) << 32) | uniform((uint32_t)co0);
            RawResult res{0, 0, 0};

            if constexpr (ALGO == ALGO_LEV || ALGO == ALGO_OSA || ALGO == ALGO_LCS) {
                // state as 32-bit halves (h* unused for 32-bit words)
                uint32_t VPl = ~0u, VPh = ~0u, VNl = 0, VNh = 0, D0l = 0, D0h = 0, PMl = 0, PMh = 0;
                const W* __restrict__ col = pm + lane;
                for (uint32_t k0 = 0; k0 < n; k0 += 64) {
                    uint32_t row = 0;
                    if (k0 + lane < n) row = look((uint32_t)c[co + k0 + lane]);
                    const uint32_t cnt = uniform(min(64u, n - k0));
                    W Xn = col[readlane(row, 0) * 64];
                    for (uint32_t i = 0; i < cnt; ++i) {
                        const W X = Xn;
                        Xn = col[readlane(row, min(i + 1, 63u)) * 64]; // prefetch the next row
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
                    }
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
                // Jaro word path; the host routes only choices with n <= 64 here.
                uint32_t row = 0;
                if (lane < n) row = look((uint32_t)c[co + lane]);
                uint32_t Bound = max(m, n) / 2;
                if (Bound > 0) Bound--;
                W BoundMask = (Bound + 1 >= BITS) ? ~(W)0 : (((W)1 << (Bound + 1)) - 1);
                W Pf = 0;
                uint64_t Tf = 0;
                uint32_t pfx = 0;
                const uint32_t maxpfx = min(min(m, n), 4u);
                for (uint32_t j = 0; j < n; ++j) {
                    const uint32_t r = readlane(row, j);
                    const W PMr = pm[r * 64 + lane];
                    if (j < maxpfx && pfx == j && ((PMr >> j) & 1)) pfx = j + 1;
                    const W x = PMr & BoundMask & ~Pf;
                    Pf |= x & ((W)0 - x);
                    Tf |= (uint64_t)(x != 0) << j;
                    BoundMask = (j < Bound) ? (W)((BoundMask << 1) | 1) : (W)(BoundMask << 1);
                }
                // count_transpositions_word (Jaro_impl.hpp:258-276), walked in text order
                uint32_t tr = 0;
                W P2 = Pf;
                for (uint32_t j = 0; j < n; ++j) {
                    const uint32_t r = readlane(row, j);
                    const W PMr = pm[r * 64 + lane];
                    if ((Tf >> j) & 1) {
                        const W pmask = P2 & ((W)0 - P2);
                        tr += (PMr & pmask) == 0;
                        P2 ^= pmask;
                    }
                }
                res.v = (uint64_t)popc<W>(Pf);
                res.tr = tr;
                res.prefix = pfx;
            }
            uint64_t bits;
            if constexpr (INTK)
                bits = finalize_int(spec, m, n, res.v);
            else
                bits = finalize_bits(spec, m, n, res);
            if (direct) {
                if (qact) ((uint64_t*)a.out)[(uint64_t)qidx * a.out_stride + cidx] = bits;
            }
            else {
                outv[t] = bits;
            }
        }
        if (!direct && qact) {
            uint64_t* dst = (uint64_t*)a.out + (uint64_t)qidx * a.out_stride + g0;
            if (gn == G && ((uintptr_t)dst & 15) == 0) {
                uint4* d4 = (uint4*)dst;
#pragma unroll
                for (uint32_t t = 0; t < G; t += 2)
                    d4[t / 2] = make_uint4((uint32_t)outv[t], (uint32_t)(outv[t] >> 32), (uint32_t)outv[t + 1],
                                           (uint32_t)(outv[t + 1] >> 32));
            }
            else if (gn == G) {
#pragma unroll
                for (uint32_t t = 0; t < G; ++t) dst[t] = outv[t];
            }
            else {
                for (uint32_t t = 0; t < gn; ++t) dst[t] = outv[t];
            }
        }
    }
}

template <int ALGO, typename CharT, typename W>
static void launch_one(const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x, size_t lds, hipStream_t s)
{
    dim3 grid(nblocks_x, ntiles);
    const bool intk = (a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY) && a.spec.metric <= RFA_OSA;
    if constexpr (ALGO != ALGO_JARO) {
        if (intk) {
            hipLaunchKernelGGL((k_tile<ALGO, CharT, W, true>), grid, dim3(TILE_THREADS), lds, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_tile<ALGO, CharT, W, false>), grid, dim3(TILE_THREADS), lds, s, a);
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
    size_t extra;
    if (char_width == 1)
        extra = 256 * 4 + 256 + 4 * TILE_WAVES;
    else
        extra = ((size_t)1 << hash_bits) * 10 + 16 + 4 * TILE_WAVES;
    return pm + extra;
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
