// rfa_editops.hip -- levenshtein_editops and lcs_seq_editops / indel_editops on
// the device (Levenshtein_impl.hpp:959-1015 and 1229-1257, LCSseq_impl.hpp:440-534).
//
// Two kernels per batch of pairs (a[i], b[i]):
//   k_ed_rows: one wave per pair.  The common prefix and suffix are found 64 units
//     at a time by ballot; lane l then holds the stripped s1[l] (at most 64 units),
//     so the match mask of a text unit is one ballot and the single-word
//     recurrence runs on wave-uniform 64-bit values:
//       Levenshtein (Hyyro 2003, Levenshtein_impl.hpp:259-309): VP / VN after
//         every text unit;
//       LCS (Hyyro 2004, LCSseq_impl.hpp:205-250): S after every text unit.
//     Text units arrive 64 at a time (one per lane, read back with v_readlane);
//     lane j keeps row j of the chunk and the chunk's rows are stored together.
//     The pair's rows start where its second string starts in b's offsets (times
//     two for VP / VN), so no host-side layout is needed.  Per pair it writes the
//     operation count, the stripped lengths and the prefix length.
//   k_ed_walk: one lane per pair.  From the bottom-right corner it takes, at
//     every cell, the operation the reference prefers, so the script is the
//     reference's:
//       Levenshtein: a deletion where VP marks the cell, else one row up an
//         insertion where VN marks it, else a diagonal (a replacement if the
//         units differ);
//       LCS: a deletion where S marks the cell (no LCS step), else one row up an
//         insertion where S is clear, else a diagonal match;
//     then the rest of either string as deletions, then insertions.  Operations
//     are written at the pair's offset (the prefix sum of the counts).
// Levenshtein's script is the reference's when the stripped first string has at
// most 64 units: the reference then records the full single-word matrix, which
// depends only on the strings.  Longer ones (up to 4096 units, where the reference
// records its block band rather than splitting by Hirschberg) go to k_ed_levl below;
// the rest are flagged.  LCS records the full matrix at every length; stripped first
// strings of up to 4096 units run here, over several words (lane l holding
// s1[64 k + l], so a word's match mask is one ballot; the words' additions carry
// across lanes by carry-lookahead on two ballots).
#include <hipcub/hipcub.hpp>

#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

size_t editops_scan_bytes(uint64_t n)
{
    size_t b = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)(n + 1));
    return b;
}

hipError_t editops_offsets(const uint64_t* counts, uint64_t* offsets, uint64_t n, void* temp, size_t bytes,
                           hipStream_t s)
{
    size_t b = bytes;
    return hipcub::DeviceScan::ExclusiveSum(temp, b, counts, offsets, (int)(n + 1), s);
}

template <typename CharT, bool LEV>
__global__ __launch_bounds__(256) void k_ed_rows(EditopsArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    for (uint64_t it = wave0; it < a.n; it += nwaves) {
        const uint64_t a0 = a.aoff[it], a1 = a.aoff[it + 1], b0 = a.boff[it], b1 = a.boff[it + 1];
        const uint64_t la0 = a1 - a0, lb0 = b1 - b0;
        const uint64_t common = la0 < lb0 ? la0 : lb0;
        // common prefix
        uint64_t pre = 0;
        for (;;) {
            const uint64_t k = pre + lane;
            const bool eq = k < common && A[a0 + k] == B[b0 + k];
            const uint64_t m = ballot(eq);
            if (m == ~0ull) {
                pre += 64;
                continue;
            }
            pre += (uint64_t)__builtin_ctzll(~m);
            break;
        }
        // common suffix of the rest
        const uint64_t rest = common - pre;
        uint64_t suf = 0;
        for (;;) {
            const uint64_t k = suf + lane;
            const bool eq = k < rest && A[a1 - 1 - k] == B[b1 - 1 - k];
            const uint64_t m = ballot(eq);
            if (m == ~0ull) {
                suf += 64;
                continue;
            }
            suf += (uint64_t)__builtin_ctzll(~m);
            break;
        }
        const uint64_t la = la0 - pre - suf, lb = lb0 - pre - suf;
        EdInfo info;
        info.pre = pre;
        info.la = (uint32_t)la;
        info.lb = (uint32_t)lb;
        info.flag = 0;
        if constexpr (LEV) {
            // past 64 units the reference records its block band, unless it splits the
            // problem (Hirschberg) first: those pairs are not taken
            if (la > 64 && lb > 0) {
                const bool whole = (2 * la * lb) / 8 < 1048576 || lb < 10;
                info.flag = la > EDITOPS_LEV_BLOCK_MAX ? 1u : (whole ? 2u : 3u);
            }
        }
        else {
            info.flag = la > EDITOPS_LCS_MAX ? 1u : 0u;
        }
        info.count = (uint32_t)(la + lb);
        if (la == 0 || lb == 0 || info.flag) {
            if (lane == 0) {
                a.info[it] = info;
                a.counts[it] = info.flag ? 0 : info.count;
                if (info.flag == 1) atomicOr(a.err, 1u);
                if (info.flag == 2)
                    atomicMax((unsigned long long*)a.need, (unsigned long long)((2 * ((la + 63) / 64) + 1) * lb));
                if (info.flag == 3) a.hlist[atomicAdd((unsigned long long*)a.hcount, 1ull)] = (uint32_t)it;
            }
            continue;
        }
        const uint64_t as = a0 + pre, bs = b0 + pre;
        const uint64_t ro = LEV ? 2 * (b0 - a.b_base) : a.row_base[it];
        if constexpr (!LEV) {
            if (la > 64) {
                // several words: lane l holds s1[64 k + l], word k's match mask is one ballot
                // and goes to lane k; the additions carry across lanes by carry-lookahead
                const uint32_t W = (uint32_t)((la + 63) / 64);
                uint32_t su[EDITOPS_LCS_MAX / 64];
#pragma unroll
                for (uint32_t k = 0; k < EDITOPS_LCS_MAX / 64; ++k)
                    su[k] = (k < W && k * 64 + lane < la) ? (uint32_t)A[as + k * 64 + lane] : 0u;
                const bool act = lane < W;
                uint64_t Sw = ~0ull;
                for (uint64_t j0 = 0; j0 < lb; j0 += 64) {
                    const uint32_t tb = j0 + lane < lb ? (uint32_t)B[bs + j0 + lane] : 0u;
                    const uint32_t steps = (uint32_t)(lb - j0 < 64 ? lb - j0 : 64);
                    for (uint32_t s = 0; s < steps; ++s) {
                        const uint32_t c = readlane(tb, s);
                        uint64_t X = 0;
#pragma unroll
                        for (uint32_t k = 0; k < EDITOPS_LCS_MAX / 64; ++k) {
                            if (k < W) {
                                const uint64_t mk = ballot(k * 64 + lane < la && su[k] == c);
                                X = lane == k ? mk : X;
                            }
                        }
                        const uint64_t u = Sw & X;
                        uint64_t x = Sw + u;
                        const bool g = act && (x < Sw);
                        const bool p = act && !g && (x == ~0ull);
                        const uint64_t G = ballot(g), Pp = ballot(p);
                        const uint64_t sum = (G | Pp) + G;
                        x += ((sum ^ Pp) >> lane) & 1u;
                        if (act) {
                            Sw = x | (Sw - u);
                            a.rows[ro + (j0 + s) * W + lane] = Sw;
                        }
                    }
                }
                const uint64_t valid = la - 64ull * lane >= 64 ? ~0ull : ((1ull << (la - 64ull * lane)) - 1);
                uint32_t cnt = act ? (uint32_t)__popcll(~Sw & valid) : 0u;
                for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d, 64);
                info.count = (uint32_t)(la + lb - 2ull * cnt);
                if (lane == 0) {
                    a.info[it] = info;
                    a.counts[it] = info.count;
                }
                continue;
            }
        }
        const bool in = lane < la;
        const uint32_t my = in ? (uint32_t)A[as + lane] : 0u;
        uint64_t VP = ~0ull, VN = 0, S = ~0ull;
        uint32_t dist = (uint32_t)la;
        const uint64_t mask = 1ull << (la - 1);
        for (uint64_t j0 = 0; j0 < lb; j0 += 64) {
            const uint32_t tb = j0 + lane < lb ? (uint32_t)B[bs + j0 + lane] : 0u;
            const uint32_t steps = (uint32_t)(lb - j0 < 64 ? lb - j0 : 64);
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
                    a.rows[ro + 2 * (j0 + lane)] = r0;
                    a.rows[ro + 2 * (j0 + lane) + 1] = r1;
                }
                else {
                    a.rows[ro + j0 + lane] = r0;
                }
            }
        }
        if constexpr (LEV)
            info.count = dist;
        else
            info.count = (uint32_t)(la + lb - 2 * (uint64_t)__popcll(~S & (la >= 64 ? ~0ull : ((1ull << la) - 1))));
        if (lane == 0) {
            a.info[it] = info;
            a.counts[it] = info.count;
        }
    }
}

template <typename CharT, bool LEV>
__global__ __launch_bounds__(256) void k_ed_walk(EditopsArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const EdInfo info = a.info[i];
    uint64_t dist = info.count;
    if (dist == 0 || info.flag >= 2) return; // block pairs and Hirschberg pieces: k_ed_levl
    EdOp* out = a.ops + a.offsets[i];
    const uint64_t pre = info.pre;
    uint64_t col = info.la, row = info.lb;
    auto put = [&](int32_t type) {
        --dist;
        EdOp o;
        o.type = type;
        o.reserved = 0;
        o.src_pos = col + pre;
        o.dest_pos = row + pre;
        out[dist] = o;
    };
    if (info.la && info.lb) {
        const uint64_t a0 = a.aoff[i] + pre, b0 = a.boff[i] + pre;
        const uint64_t* __restrict__ r = a.rows + (LEV ? 2 * (a.boff[i] - a.b_base) : a.row_base[i]);
        const uint64_t W = (info.la + 63) / 64; // LCS row words
        while (row && col) {
            const uint64_t c = col - 1;
            const bool del = LEV ? ((r[2 * (row - 1)] >> c) & 1) : ((r[(row - 1) * W + c / 64] >> (c % 64)) & 1);
            if (del) {
                --col;
                put(RFA_EDIT_DELETE);
                continue;
            }
            --row;
            const bool ins =
                row && (LEV ? ((r[2 * (row - 1) + 1] >> c) & 1) : !((r[(row - 1) * W + c / 64] >> (c % 64)) & 1));
            if (ins) {
                put(RFA_EDIT_INSERT);
                continue;
            }
            --col;
            if (LEV && A[a0 + col] != B[b0 + row]) put(RFA_EDIT_REPLACE);
        }
    }
    while (col) {
        --col;
        put(RFA_EDIT_DELETE);
    }
    while (row) {
        --row;
        put(RFA_EDIT_INSERT);
    }
}

// ---- Levenshtein editops over several words (Levenshtein_impl.hpp:641-850, 959-1045) ----
//
// A stripped first string of 65..4096 units: the reference records the band of its
// block recurrence (max = the longer stripped length, tightened row by row), with
// Ukkonen's band of blocks [first_block, last_block] moving per row, and walks it back
// through ShiftedBitMatrix: row r keeps the words from its first block on (bits left
// of them read as clear), words the row did not advance keep the matrix's initial
// values (VP all set, VN clear).  The same steps here: one wave per pair, word w in
// lane w.  A row advances the band's words from first_block with the carries
// HP = 1, HN = 0 entering it; a word's D0 depends on the HN carry from the word
// below it, so every lane forms D0 for both carry values, the chain of HN carries is
// resolved over the band by scalar steps on two ballots, and each lane then takes its
// D0.  The band bookkeeping (the running max, a block added after the last one, the
// shrinking and advancing ends) reads the words' scores with v_readlane.
//   RECORD = false: the pair's distance (its operation count).
//   RECORD = true: the rows into the wave's scratch, then lane 0 walks them back and
//   writes the operations at the pair's offset.

__device__ __forceinline__ uint32_t lev_row_num(uint32_t w, uint32_t W, uint32_t la)
{
    return w + 1 == W ? la - 1 : (w + 1) * 64 - 1;
}

// One stripped pair (A + as, la units; B + bs, lb units; la >= 1, lb >= 1) through the band:
// its distance, and with RECORD its rows in `rows` and its operations at out (positions
// offset by src_off / dst_off) -- only when the distance is the expected one (cap), as the
// caller sized out by it; otherwise *bad_out.
template <typename CharT, bool RECORD>
__device__ void levl_one(const CharT* __restrict__ A, uint64_t as, uint32_t la, const CharT* __restrict__ B,
                         uint64_t bs, uint32_t lb, uint64_t src_off, uint64_t dst_off, uint64_t* __restrict__ rows,
                         EdOp* __restrict__ out, uint32_t cap, uint32_t* dist_out, bool* bad_out)
{
    const uint32_t lane = threadIdx.x & 63;
    constexpr uint32_t WMAX = EDITOPS_LEV_BLOCK_MAX / 64;
    const uint32_t W = (la + 63) / 64; // 1..WMAX
    uint32_t su[WMAX];
#pragma unroll
    for (uint32_t k = 0; k < WMAX; ++k)
        su[k] = (k < W && k * 64 + lane < la) ? (uint32_t)A[as + k * 64 + lane] : 0u;
    const uint64_t Last = 1ull << ((la - 1) % 64);
    const bool top = lane == W - 1;
    uint64_t VP = ~0ull, VN = 0;
    int64_t score = lane + 1 < W ? (int64_t)(lane + 1) * 64 : (int64_t)la;
    int64_t maxv = la > lb ? la : lb;
    uint32_t fb = 0;
    uint32_t lbk;
    {
        const int64_t t = min(maxv, (maxv + (int64_t)la - (int64_t)lb) / 2) + 1;
        lbk = (uint32_t)min<int64_t>((int64_t)W, (t + 63) / 64) - 1;
    }
    bool bad = false;
    const uint32_t rw = 2 * W + 1; // row: first block, W VP words, W VN words
    for (uint32_t j0 = 0; j0 < lb && !bad; j0 += 64) {
        const uint32_t tb = j0 + lane < lb ? (uint32_t)B[bs + j0 + lane] : 0u;
        const uint32_t steps = min(64u, lb - j0);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t row = j0 + s;
            const uint32_t c = readlane(tb, s);
            uint64_t PM = 0; // lane k: the match mask of word k
#pragma unroll
            for (uint32_t k = 0; k < WMAX; ++k) {
                if (k < W) {
                    const uint64_t m = ballot(k * 64 + lane < la && su[k] == c);
                    PM = lane == k ? m : PM;
                }
            }
            const uint32_t fb0 = fb, lb0 = lbk;
            const bool inb = lane >= fb0 && lane <= lb0;
            // D0 for an HN carry of 0 and of 1 into the word
            const uint64_t X0 = PM, X1 = PM | 1ull;
            const uint64_t D00 = (((X0 & VP) + VP) ^ VP) | X0 | VN;
            const uint64_t D01 = (((X1 & VP) + VP) ^ VP) | X1 | VN;
            const bool h0 = top ? ((D00 & VP) & Last) != 0 : ((D00 & VP) >> 63) != 0;
            const bool h1 = top ? ((D01 & VP) & Last) != 0 : ((D01 & VP) >> 63) != 0;
            const uint64_t M0 = ballot(inb && h0), M1 = ballot(inb && h1);
            uint64_t H = 0; // bit w: the HN carry into word w
            {
                uint32_t h = 0;
                for (uint32_t w = fb0; w <= lb0; ++w) {
                    H |= (uint64_t)h << w;
                    h = (uint32_t)(((h ? M1 : M0) >> w) & 1u);
                }
            }
            const uint64_t hn_in = (H >> lane) & 1u;
            const uint64_t D0 = hn_in ? D01 : D00;
            uint64_t HP = VN | ~(D0 | VP);
            uint64_t HN = D0 & VP;
            const bool hpo = top ? (HP & Last) != 0 : (HP >> 63) != 0;
            const bool hno = top ? (HN & Last) != 0 : (HN >> 63) != 0;
            const uint64_t PO = ballot(inb && hpo), NO = ballot(inb && hno);
            const uint64_t hp_in = lane == fb0 ? 1ull : (lane ? (PO >> (lane - 1)) & 1u : 0ull);
            HP = (HP << 1) | hp_in;
            HN = (HN << 1) | hn_in;
            if (inb) {
                VP = HN | ~(D0 | HP);
                VN = HP & D0;
                score += (int64_t)hpo - (int64_t)hno;
            }
            // the running bound on the distance
            {
                const int64_t sl = (int64_t)(int32_t)readlane((uint32_t)score, lb0);
                const int64_t r1 = (int64_t)lb - (int64_t)row - 1;
                const int64_t r2 = (int64_t)la - ((int64_t)(1 + lb0) * 64 - 2);
                maxv = min(maxv, sl + max(r1, r2));
            }
            uint32_t amax = lb0; // the last word this row advanced
            if (lb0 + 1 < W) {
                const int64_t sl = (int64_t)(int32_t)readlane((uint32_t)score, lb0);
                const int64_t cond = (maxv + 128 + row + la) - (sl + 2 + lb);
                if ((int64_t)lev_row_num(lb0, W, la) < cond) {
                    // the next block joins with the carries out of the last one
                    const uint32_t nw = lb0 + 1;
                    const uint64_t hpc = (PO >> lb0) & 1u, hnc = (NO >> lb0) & 1u;
                    if (lane == nw) {
                        const int64_t chars = nw + 1 == W ? (int64_t)((la - 1) % 64 + 1) : 64;
                        score = sl + chars - (int64_t)hpc + (int64_t)hnc;
                        VP = ~0ull;
                        VN = 0;
                        const uint64_t X = PM | hnc;
                        const uint64_t D = (((X & VP) + VP) ^ VP) | X | VN;
                        uint64_t hp = VN | ~(D | VP);
                        uint64_t hn = D & VP;
                        const bool po = top ? (hp & Last) != 0 : (hp >> 63) != 0;
                        const bool no = top ? (hn & Last) != 0 : (hn >> 63) != 0;
                        hp = (hp << 1) | hpc;
                        hn = (hn << 1) | hnc;
                        VP = hn | ~(D | hp);
                        VN = hp & D;
                        score += (int64_t)po - (int64_t)no;
                    }
                    amax = nw;
                    lbk = nw;
                }
            }
            if constexpr (RECORD) {
                // row: the first block, then VP and VN of the words from it on
                uint64_t* r = rows + (uint64_t)row * rw;
                const uint32_t w = fb0 + lane;
                const bool have = w <= amax;
                const uint32_t src = have ? w : lane;
                const uint64_t vp = __shfl(VP, (int)src, 64), vn = __shfl(VN, (int)src, 64);
                if (lane < W) {
                    r[1 + lane] = have ? vp : ~0ull;
                    r[1 + W + lane] = have ? vn : 0ull;
                }
                if (lane == 0) r[0] = fb0;
            }
            // Ukkonen's band for the next row
            for (;;) {
                const int64_t sl = (int64_t)(int32_t)readlane((uint32_t)score, lbk);
                const bool c1 = sl < maxv + 64;
                const int64_t cond = (maxv + 128 + row + la + 1) - (sl + 2 + lb);
                const bool c2 = (int64_t)lev_row_num(lbk, W, la) <= cond;
                if (c1 && c2) break;
                if (lbk == fb) {
                    bad = true;
                    break;
                }
                --lbk;
            }
            if (bad) break;
            for (; fb <= lbk; ++fb) {
                const int64_t sf = (int64_t)(int32_t)readlane((uint32_t)score, fb);
                const bool c1 = sf < maxv + 64;
                const int64_t cond = (sf + la + row) - (maxv + lb);
                const bool c2 = (int64_t)lev_row_num(fb, W, la) >= cond;
                if (c1 && c2) break;
            }
            if (lbk < fb) {
                bad = true;
                break;
            }
        }
    }
    const uint32_t dist = readlane((uint32_t)score, W - 1);
    *dist_out = dist;
    if (RECORD && dist != cap) bad = true;
    *bad_out = bad;
    if constexpr (RECORD) {
        __threadfence_block();
        if (lane == 0 && !bad && dist) {
            uint64_t d = dist, col = la, row = lb;
            auto put = [&](int32_t type) {
                --d;
                EdOp o;
                o.type = type;
                o.reserved = 0;
                o.src_pos = col + src_off;
                o.dest_pos = row + dst_off;
                out[d] = o;
            };
            // ShiftedBitMatrix::test_bit: clear left of the row's first block
            auto bit = [&](uint64_t r, uint64_t cc, uint32_t which) -> bool {
                const uint64_t* rr = rows + r * rw;
                const uint64_t f = rr[0], wd = cc / 64;
                if (wd < f) return false;
                return (rr[1 + which * W + (wd - f)] >> (cc % 64)) & 1u;
            };
            while (row && col) {
                if (bit(row - 1, col - 1, 0)) {
                    --col;
                    put(RFA_EDIT_DELETE);
                    continue;
                }
                --row;
                if (row && bit(row - 1, col - 1, 1)) {
                    put(RFA_EDIT_INSERT);
                    continue;
                }
                --col;
                if (A[as + col] != B[bs + row]) put(RFA_EDIT_REPLACE);
            }
            while (col) {
                --col;
                put(RFA_EDIT_DELETE);
            }
            while (row) {
                --row;
                put(RFA_EDIT_INSERT);
            }
        }
    }
}

template <typename CharT, bool RECORD>
__global__ __launch_bounds__(256) void k_ed_levl(EditopsArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = (uint64_t)blockIdx.x * 4 + uniform(threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    uint64_t* __restrict__ rows = RECORD ? a.long_rows + wave0 * a.long_stride : nullptr;
    if (a.leaves) {
        // pieces of the pairs split by Hirschberg (RECORD only)
        for (uint64_t it = wave0; it < a.nleaves; it += nwaves) {
            const EdLeaf L = a.leaves[it];
            const uint32_t la = uniform(L.la), lb = uniform(L.lb);
            const uint64_t as = a.aoff[L.pair] + L.a0, bs = a.boff[L.pair] + L.b0;
            EdOp* out = a.ops + L.opos;
            if (la == 0 || lb == 0) { // the rest of either string (recover_alignment's tails)
                for (uint32_t i = lane; i < la; i += 64) {
                    EdOp o;
                    o.type = RFA_EDIT_DELETE;
                    o.reserved = 0;
                    o.src_pos = L.src + i;
                    o.dest_pos = L.dst;
                    out[i] = o;
                }
                for (uint32_t j = lane; j < lb; j += 64) {
                    EdOp o;
                    o.type = RFA_EDIT_INSERT;
                    o.reserved = 0;
                    o.src_pos = L.src;
                    o.dest_pos = L.dst + j;
                    out[j] = o;
                }
                continue;
            }
            uint32_t dist;
            bool bad;
            levl_one<CharT, true>(A, as, la, B, bs, lb, L.src, L.dst, rows, out, uniform(L.dist), &dist, &bad);
            if (bad && lane == 0) atomicOr(a.err, 1u);
        }
        return;
    }
    // the pairs round robin over the waves: wave w, lane l of chunk k looks at w + (64 k + l) * nwaves
    for (uint64_t base = 0; wave0 + base * nwaves < a.n; base += 64) {
        const uint64_t idx = wave0 + (base + lane) * nwaves;
        // 2: walked here; 3 (Hirschberg): counted here, walked as leaves
        uint64_t todo = ballot(idx < a.n && (a.info[idx].flag == 2 || (!RECORD && a.info[idx].flag == 3)));
        while (todo) {
            const uint64_t it = wave0 + (base + (uint64_t)__builtin_ctzll(todo)) * nwaves;
            todo &= todo - 1;
            const uint64_t pre = a.info[it].pre;
            const uint32_t la = uniform(a.info[it].la), lb = uniform(a.info[it].lb);
            const uint64_t as = a.aoff[it] + pre, bs = a.boff[it] + pre;
            uint32_t dist;
            bool bad;
            const uint32_t cap = RECORD ? uniform((uint32_t)(a.offsets[it + 1] - a.offsets[it])) : 0u;
            levl_one<CharT, RECORD>(A, as, la, B, bs, lb, pre, pre, rows, RECORD ? a.ops + a.offsets[it] : nullptr,
                                    cap, &dist, &bad);
            if (lane == 0) {
                if (bad) {
                    a.info[it].flag = 1;
                    atomicOr(a.err, 1u);
                }
                else if (!RECORD) {
                    a.info[it].count = dist;
                    a.counts[it] = dist;
                }
            }
        }
    }
}

// find_hirschberg_pos (Levenshtein_impl.hpp:1098-1182) of a piece: s2 split at lb / 2; the
// forward row D(s1[0, i), s2[0, lb/2)) and the backward row D(s1[i, la), s2[lb/2, lb)) for
// every i, from the block recurrence over all of s1's words (lane w: word w; no band), then
// the first i of the smallest sum.  tmp: the two rows (la + 1 values each) of the wave.
template <typename CharT>
__device__ void split_rows(const CharT* __restrict__ A, uint64_t as, uint32_t la, const CharT* __restrict__ B,
                           uint64_t bs, uint32_t t0, uint32_t tn, bool rev, uint32_t* __restrict__ D)
{
    const uint32_t lane = threadIdx.x & 63;
    constexpr uint32_t WMAX = EDITOPS_LEV_BLOCK_MAX / 64;
    const uint32_t W = (la + 63) / 64;
    uint32_t su[WMAX]; // lane l: s1[64 k + l], or the reversed s1's
#pragma unroll
    for (uint32_t k = 0; k < WMAX; ++k) {
        const uint32_t i = k * 64 + lane;
        su[k] = (k < W && i < la) ? (uint32_t)A[as + (rev ? la - 1 - i : i)] : 0u;
    }
    const uint64_t Last = 1ull << ((la - 1) % 64);
    const bool top = lane == W - 1, act = lane < W;
    uint64_t VP = ~0ull, VN = 0;
    for (uint32_t j0 = 0; j0 < tn; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint32_t tb = j < tn ? (uint32_t)B[bs + (rev ? t0 + tn - 1 - j : t0 + j)] : 0u;
        const uint32_t steps = min(64u, tn - j0);
        for (uint32_t s = 0; s < steps; ++s) {
            const uint32_t c = readlane(tb, s);
            uint64_t PM = 0;
#pragma unroll
            for (uint32_t k = 0; k < WMAX; ++k) {
                if (k < W) {
                    const uint64_t m = ballot(k * 64 + lane < la && su[k] == c);
                    PM = lane == k ? m : PM;
                }
            }
            const uint64_t X0 = PM, X1 = PM | 1ull;
            const uint64_t D00 = (((X0 & VP) + VP) ^ VP) | X0 | VN;
            const uint64_t D01 = (((X1 & VP) + VP) ^ VP) | X1 | VN;
            const bool h0 = top ? ((D00 & VP) & Last) != 0 : ((D00 & VP) >> 63) != 0;
            const bool h1 = top ? ((D01 & VP) & Last) != 0 : ((D01 & VP) >> 63) != 0;
            const uint64_t M0 = ballot(act && h0), M1 = ballot(act && h1);
            uint64_t H = 0;
            {
                uint32_t h = 0;
                for (uint32_t w = 0; w < W; ++w) {
                    H |= (uint64_t)h << w;
                    h = (uint32_t)(((h ? M1 : M0) >> w) & 1u);
                }
            }
            const uint64_t hn_in = (H >> lane) & 1u;
            const uint64_t D0 = hn_in ? D01 : D00;
            uint64_t HP = VN | ~(D0 | VP);
            uint64_t HN = D0 & VP;
            const bool hpo = top ? (HP & Last) != 0 : (HP >> 63) != 0;
            const uint64_t PO = ballot(act && hpo);
            const uint64_t hp_in = lane == 0 ? 1ull : ((PO >> (lane - 1)) & 1u);
            HP = (HP << 1) | hp_in;
            HN = (HN << 1) | hn_in;
            if (act) {
                VP = HN | ~(D0 | HP);
                VN = HP & D0;
            }
        }
    }
    // D(s1[0, i), text) = tn + the vertical steps above row i
    const uint64_t valid = !act ? 0ull : (la - 64ull * lane >= 64 ? ~0ull : ((1ull << (la - 64ull * lane)) - 1));
    const int32_t wsum = (int32_t)__popcll(VP & valid) - (int32_t)__popcll(VN & valid);
    int32_t x = wsum; // inclusive scan over the words
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(x, d, 64);
        if ((int)lane >= d) x += y;
    }
    int32_t base = (int32_t)tn + (x - wsum);
    if (lane == 0) D[0] = tn;
    if (act) {
        for (uint32_t b = 0; b < 64 && 64 * lane + b < la; ++b) {
            base += (int32_t)((VP >> b) & 1u) - (int32_t)((VN >> b) & 1u);
            D[64 * lane + b + 1] = (uint32_t)base;
        }
    }
}

template <typename CharT>
__global__ __launch_bounds__(64) void k_ed_split(EditopsArgs a, EdSplit* __restrict__ splits, uint64_t nsplits,
                                                 uint32_t* __restrict__ tmp, uint64_t tmp_stride)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t* L = tmp + (uint64_t)blockIdx.x * tmp_stride;
    for (uint64_t it = blockIdx.x; it < nsplits; it += gridDim.x) {
        const EdSplit sp = splits[it];
        const uint32_t la = uniform(sp.la), lb = uniform(sp.lb);
        const uint64_t as = a.aoff[sp.pair] + sp.a0, bs = a.boff[sp.pair] + sp.b0;
        const uint32_t ls = lb / 2;
        uint32_t* R = L + la + 1;
        split_rows<CharT>(A, as, la, B, bs, 0, ls, false, L);
        split_rows<CharT>(A, as, la, B, bs, ls, lb - ls, true, R); // R[j] = D(s1[la - j, la), s2[ls, lb))
        __syncthreads();
        uint64_t best = ~0ull; // (sum << 32) | i: the smallest sum, then the first i
        for (uint32_t i = lane; i <= la; i += 64) {
            const uint64_t key = ((uint64_t)(L[i] + R[la - i]) << 32) | i;
            best = key < best ? key : best;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t o = __shfl_xor(best, off, 64);
            best = o < best ? o : best;
        }
        if (lane == 0) {
            const uint32_t mid = (uint32_t)best;
            splits[it].mid = mid;
            splits[it].left = L[mid];
            splits[it].right = R[la - mid];
        }
        __syncthreads();
    }
}

hipError_t launch_editops_split(int char_width, const EditopsArgs& a, EdSplit* splits, uint64_t nsplits,
                                uint32_t* tmp, uint64_t tmp_stride, uint32_t nblocks, hipStream_t s)
{
    if (nsplits == 0) return hipSuccess;
    if (char_width == 1)
        hipLaunchKernelGGL((k_ed_split<uint8_t>), dim3(nblocks), dim3(64), 0, s, a, splits, nsplits, tmp, tmp_stride);
    else if (char_width == 2)
        hipLaunchKernelGGL((k_ed_split<uint16_t>), dim3(nblocks), dim3(64), 0, s, a, splits, nsplits, tmp, tmp_stride);
    else
        hipLaunchKernelGGL((k_ed_split<uint32_t>), dim3(nblocks), dim3(64), 0, s, a, splits, nsplits, tmp, tmp_stride);
    return hipGetLastError();
}

template <typename CharT>
static void launch_ed_levl_c(const EditopsArgs& a, bool walk, uint32_t nblocks, hipStream_t s)
{
    if (walk)
        hipLaunchKernelGGL((k_ed_levl<CharT, true>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_ed_levl<CharT, false>), dim3(nblocks), dim3(256), 0, s, a);
}

hipError_t launch_editops_lev_block(int char_width, const EditopsArgs& a, bool walk, uint32_t nblocks,
                                    hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    if (char_width == 1)
        launch_ed_levl_c<uint8_t>(a, walk, nblocks, s);
    else if (char_width == 2)
        launch_ed_levl_c<uint16_t>(a, walk, nblocks, s);
    else
        launch_ed_levl_c<uint32_t>(a, walk, nblocks, s);
    return hipGetLastError();
}

// hamming_editops with pad (Hamming_impl.hpp:40-59): a replacement at every differing
// position of the common length, then deletions of s1's tail or insertions of s2's.
// One wave per pair; the count is a ballot popcount per 64 positions, and the write
// pass ranks each differing position among the wave's by mbcnt, so the operations are
// written in parallel and in order.
template <typename CharT, bool WRITE>
__global__ __launch_bounds__(256) void k_ed_hamming(EditopsArgs a)
{
    const CharT* __restrict__ A = (const CharT*)a.achars;
    const CharT* __restrict__ B = (const CharT*)a.bchars;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t nwaves = (uint64_t)gridDim.x * 4;
    for (uint64_t it = wave0; it < a.n; it += nwaves) {
        const uint64_t a0 = a.aoff[it], b0 = a.boff[it];
        const uint64_t la = a.aoff[it + 1] - a0, lb = a.boff[it + 1] - b0;
        const uint64_t mn = la < lb ? la : lb;
        EdOp* out = WRITE ? a.ops + a.offsets[it] : nullptr;
        uint64_t k = 0;
        for (uint64_t i0 = 0; i0 < mn; i0 += 64) {
            const uint64_t i = i0 + lane;
            const bool diff = i < mn && A[a0 + i] != B[b0 + i];
            const uint64_t m = ballot(diff);
            if constexpr (WRITE) {
                if (diff) {
                    EdOp o;
                    o.type = RFA_EDIT_REPLACE;
                    o.reserved = 0;
                    o.src_pos = i;
                    o.dest_pos = i;
                    out[k + __popcll(m & ((1ull << lane) - 1))] = o;
                }
            }
            k += (uint64_t)__popcll(m);
        }
        if constexpr (WRITE) {
            for (uint64_t i = mn + lane; i < la; i += 64) {
                EdOp o;
                o.type = RFA_EDIT_DELETE;
                o.reserved = 0;
                o.src_pos = i;
                o.dest_pos = lb;
                out[k + (i - mn)] = o;
            }
            for (uint64_t i = mn + lane; i < lb; i += 64) {
                EdOp o;
                o.type = RFA_EDIT_INSERT;
                o.reserved = 0;
                o.src_pos = la;
                o.dest_pos = i;
                out[k + (i - mn)] = o;
            }
        }
        else if (lane == 0) {
            EdInfo info;
            info.pre = 0;
            info.la = (uint32_t)la;
            info.lb = (uint32_t)lb;
            info.flag = 0;
            info.count = (uint32_t)(k + (la - mn) + (lb - mn));
            a.info[it] = info;
            a.counts[it] = info.count;
        }
    }
}

template <typename CharT>
static void launch_ed_hamming_c(const EditopsArgs& a, bool walk, uint32_t nblocks, hipStream_t s)
{
    if (walk)
        hipLaunchKernelGGL((k_ed_hamming<CharT, true>), dim3(nblocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_ed_hamming<CharT, false>), dim3(nblocks), dim3(256), 0, s, a);
}

hipError_t launch_editops_hamming(int char_width, const EditopsArgs& a, bool walk, uint32_t nblocks, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    if (char_width == 1)
        launch_ed_hamming_c<uint8_t>(a, walk, nblocks, s);
    else if (char_width == 2)
        launch_ed_hamming_c<uint16_t>(a, walk, nblocks, s);
    else
        launch_ed_hamming_c<uint32_t>(a, walk, nblocks, s);
    return hipGetLastError();
}

template <typename CharT>
static void launch_ed_c(const EditopsArgs& a, bool lev, bool walk, uint32_t nblocks, hipStream_t s)
{
    if (walk) {
        const uint32_t nb = (uint32_t)((a.n + 255) / 256);
        if (lev)
            hipLaunchKernelGGL((k_ed_walk<CharT, true>), dim3(nb), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_ed_walk<CharT, false>), dim3(nb), dim3(256), 0, s, a);
    }
    else {
        if (lev)
            hipLaunchKernelGGL((k_ed_rows<CharT, true>), dim3(nblocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_ed_rows<CharT, false>), dim3(nblocks), dim3(256), 0, s, a);
    }
}

hipError_t launch_editops(int char_width, const EditopsArgs& a, bool lev, bool walk, uint32_t nblocks,
                          hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    if (char_width == 1)
        launch_ed_c<uint8_t>(a, lev, walk, nblocks, s);
    else if (char_width == 2)
        launch_ed_c<uint16_t>(a, lev, walk, nblocks, s);
    else
        launch_ed_c<uint32_t>(a, lev, walk, nblocks, s);
    return hipGetLastError();
}

} // namespace rfa
