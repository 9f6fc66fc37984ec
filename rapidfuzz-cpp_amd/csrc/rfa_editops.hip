// rfa_editops.hip -- the bit matrices behind levenshtein_editops and
// lcs_seq_editops / indel_editops (Levenshtein_impl.hpp:1229-1257,
// LCSseq_impl.hpp:528-534).  The host strips each pair's common affix and walks
// the recorded rows back into edit operations; the device records the rows.
//
// One wave per pair: lane l holds s1[l] (the stripped first string, at most 64
// units), so the match mask of a text unit is one ballot, and the single-word
// recurrences run on wave-uniform 64-bit values:
//   Levenshtein (Hyyro 2003, Levenshtein_impl.hpp:259-309): VP / VN after every
//     text unit, plus the final distance;
//   LCS (Hyyro 2004, LCSseq_impl.hpp:205-250): S after every text unit, plus the
//     LCS length.
// Text units arrive 64 at a time (one per lane, read back with v_readlane); lane
// j keeps row j of the chunk and the chunk's rows are stored together.
#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

template <typename CharT, bool LEV>
__global__ __launch_bounds__(256) void k_editops(EditopsArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave0 = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t nwaves = gridDim.x * 4;
    for (uint32_t it = wave0; it < a.n; it += nwaves) {
        const uint32_t la = uniform(a.a_len[it]), lb = uniform(a.b_len[it]);
        const uint64_t as = a.a_start[it], bs = a.b_start[it], ro = a.row_off[it];
        const bool in = lane < la;
        const uint32_t my = in ? (uint32_t)A[as + lane] : 0u;
        uint64_t VP = ~0ull, VN = 0, S = ~0ull;
        uint32_t dist = la;
        const uint64_t mask = la ? 1ull << (la - 1) : 0;
        for (uint32_t j0 = 0; j0 < lb; j0 += 64) {
            const uint32_t tb = j0 + lane < lb ? (uint32_t)B[bs + j0 + lane] : 0u;
            const uint32_t steps = min(64u, lb - j0);
            uint64_t r0 = 0, r1 = 0;
            for (uint32_t s = 0; s < steps; ++s) {
                const uint32_t c = readlane(tb, s);
                const uint64_t M = ballot(in && my == c);
                if constexpr (LEV) {
                    const uint64_t X = M | VN;
                    const uint64_t D0 = (((X & VP) + VP) ^ VP) | X;
                    uint64_t HP = VN | ~(D0 | VP);
                    uint64_t HN = D0 & VP;
                    dist += (HP & mask) != 0;
                    dist -= (HN & mask) != 0;
                    HP = (HP << 1) | 1;
                    HN <<= 1;
                    VP = HN | ~(D0 | HP);
                    VN = HP & D0;
                    if (lane == s) {
                        r0 = VP;
                        r1 = VN;
                    }
                }
                else {
                    const uint64_t u = S & M;
                    S = (S + u) | (S - u);
                    if (lane == s) r0 = S;
                }
            }
            if (lane < steps) {
                if constexpr (LEV) {
                    a.rows[ro + 2 * (uint64_t)(j0 + lane)] = r0;
                    a.rows[ro + 2 * (uint64_t)(j0 + lane) + 1] = r1;
                }
                else {
                    a.rows[ro + j0 + lane] = r0;
                }
            }
        }
        if (lane == 0) {
            if constexpr (LEV)
                a.score[it] = dist;
            else
                a.score[it] = (uint32_t)__popcll(~S & (la >= 64 ? ~0ull : ((1ull << la) - 1)));
        }
    }
}

template <typename CharT>
static void launch_editops_c(const EditopsArgs& a, bool lev, uint32_t nblocks, hipStream_t s)
{
    if (lev)
        hipLaunchKernelGGL((k_editops<CharT, true>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_editops<CharT, false>), dim3(nblocks), dim3(256), 0, s, a);
}

hipError_t launch_editops(int char_width, const EditopsArgs& a, bool lev, uint32_t nblocks, hipStream_t s)
{
    if (char_width == 1)
        launch_editops_c<uint8_t>(a, lev, nblocks, s);
    else if (char_width == 2)
        launch_editops_c<uint16_t>(a, lev, nblocks, s);
    else
        launch_editops_c<uint32_t>(a, lev, nblocks, s);
    return hipGetLastError();
}

} // namespace rfa
