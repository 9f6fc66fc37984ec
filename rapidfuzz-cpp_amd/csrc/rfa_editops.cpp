// rfa_editops.cpp -- rfa_editops: levenshtein_editops and lcs_seq_editops /
// indel_editops for batches of pairs (Levenshtein_impl.hpp:959-1015 and 1229-1257,
// LCSseq_impl.hpp:440-534).
//
// Per pair the host strips the common prefix and suffix (they add no operation),
// the device records the bit matrix of the rest (rfa_editops.hip), and the host
// walks it back from the bottom-right corner, preferring at every cell the
// operation the reference prefers, so the scripts are the reference's:
//   Levenshtein: a deletion where VP marks the cell, else step up a row and take
//     an insertion where VN marks it, else a diagonal (a replacement if the units
//     differ);
//   LCS: a deletion where S marks the cell (no LCS step), else step up a row and
//     take an insertion where S is clear there, else a diagonal match;
//   whatever is left of either string at the edge becomes deletions, then
//   insertions.
// Levenshtein's walk is the reference's when the stripped first string has at
// most 64 units: the reference then records the full single-word matrix, which
// depends only on the strings.  (Longer ones use banded or Hirschberg matrices.)
#include <algorithm>
#include <cstring>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_internal.hpp"

using namespace rfa;

namespace {

uint32_t unit_at(const rfa_strings* s, uint64_t u)
{
    const uint8_t* p = s->h_chars.data() + u * (uint64_t)s->width;
    switch (s->width) {
    case 1: return p[0];
    case 2: {
        uint16_t v;
        std::memcpy(&v, p, 2);
        return v;
    }
    default: {
        uint32_t v;
        std::memcpy(&v, p, 4);
        return v;
    }
    }
}

struct EdPair {
    uint64_t a0 = 0, b0 = 0; // first unit of each stripped string (absolute unit index)
    uint32_t la = 0, lb = 0; // stripped lengths
    uint64_t pre = 0;        // common prefix length
    int64_t item = -1;       // device item, -1 when a stripped string is empty
    uint64_t row_off = 0;    // first row word of the item
};

inline bool bit(uint64_t w, uint64_t i) { return (w >> i) & 1u; }

// Walks one pair's rows back into dist operations at out[0, dist).
void walk(const EdPair& p, bool lev, const uint64_t* rows, uint64_t dist, const rfa_strings* a,
          const rfa_strings* b, rfa_editop* out)
{
    uint64_t col = p.la, row = p.lb;
    auto put = [&](int32_t type) {
        --dist;
        out[dist].type = type;
        out[dist].reserved = 0;
        out[dist].src_pos = col + p.pre;
        out[dist].dest_pos = row + p.pre;
    };
    if (p.item >= 0) {
        const uint64_t* r = rows + p.row_off;
        while (row && col) {
            const uint64_t c = col - 1;
            const bool del = lev ? bit(r[2 * (row - 1)], c) : bit(r[row - 1], c);
            if (del) {
                --col;
                put(RFA_EDIT_DELETE);
                continue;
            }
            --row;
            const bool ins = row && (lev ? bit(r[2 * (row - 1) + 1], c) : !bit(r[row - 1], c));
            if (ins) {
                put(RFA_EDIT_INSERT);
                continue;
            }
            --col;
            if (lev && unit_at(a, p.a0 + col) != unit_at(b, p.b0 + row)) put(RFA_EDIT_REPLACE);
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

void editops_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, int metric, uint64_t* offsets,
                  rfa_editop* ops, uint64_t capacity)
{
    if (metric != RFA_LEVENSHTEIN && metric != RFA_INDEL && metric != RFA_LCSSEQ)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "editops are for Levenshtein, Indel and LCSseq");
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    if (a->h_chars.size() < a->h_off.back() * a->width || b->h_chars.size() < b->h_off.back() * b->width)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "editops need string sets that hold their host copy");
    const bool lev = metric == RFA_LEVENSHTEIN;
    const uint64_t n = a->count;
    std::vector<EdPair> pairs(n);
    std::vector<uint64_t> it_a0, it_b0, it_row;
    std::vector<uint32_t> it_la, it_lb;
    uint64_t row_words = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t as = a->h_off[i], ae = a->h_off[i + 1], bs = b->h_off[i], be = b->h_off[i + 1];
        uint64_t pre = 0;
        while (as + pre < ae && bs + pre < be && unit_at(a, as + pre) == unit_at(b, bs + pre)) ++pre;
        as += pre;
        bs += pre;
        while (as < ae && bs < be && unit_at(a, ae - 1) == unit_at(b, be - 1)) {
            --ae;
            --be;
        }
        if (ae - as > 64)
            throw ApiError(RFA_ERR_UNSUPPORTED, "editops of a pair whose first string has more than 64 code units "
                                                "without the common prefix and suffix");
        EdPair& p = pairs[i];
        p.a0 = as;
        p.b0 = bs;
        p.la = (uint32_t)(ae - as);
        p.lb = (uint32_t)(be - bs);
        p.pre = pre;
        if (p.la && p.lb) {
            p.item = (int64_t)it_a0.size();
            p.row_off = row_words;
            it_a0.push_back(as);
            it_b0.push_back(bs);
            it_la.push_back(p.la);
            it_lb.push_back(p.lb);
            it_row.push_back(row_words);
            row_words += (lev ? 2ull : 1ull) * p.lb;
        }
    }
    if (it_a0.size() > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 - 1 pairs");
    const uint32_t items = (uint32_t)it_a0.size();
    std::vector<uint32_t> score(items);
    std::vector<uint64_t> rows(row_words);
    if (items) {
        // items: a_start, b_start, row_off (u64) then a_len, b_len, score (u32)
        const size_t bytes = (size_t)items * (3 * 8 + 3 * 4);
        uint8_t* d = (uint8_t*)ctx->ed_items.get(bytes);
        hipStream_t s = ctx->stream;
        EditopsArgs ea{};
        ea.achars = a->d_chars;
        ea.bchars = b->d_chars;
        ea.a_start = (const uint64_t*)d;
        ea.b_start = (const uint64_t*)(d + (size_t)items * 8);
        ea.row_off = (const uint64_t*)(d + (size_t)items * 16);
        ea.a_len = (const uint32_t*)(d + (size_t)items * 24);
        ea.b_len = (const uint32_t*)(d + (size_t)items * 28);
        ea.score = (uint32_t*)(d + (size_t)items * 32);
        ea.rows = (uint64_t*)ctx->ed_rows.get(std::max<uint64_t>(8, row_words * 8));
        ea.n = items;
        HIP_CHECK(hipMemcpyAsync((void*)ea.a_start, it_a0.data(), (size_t)items * 8, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((void*)ea.b_start, it_b0.data(), (size_t)items * 8, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((void*)ea.row_off, it_row.data(), (size_t)items * 8, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((void*)ea.a_len, it_la.data(), (size_t)items * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync((void*)ea.b_len, it_lb.data(), (size_t)items * 4, hipMemcpyHostToDevice, s));
        const uint32_t nblocks =
            (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((items + 3) / 4, (uint64_t)ctx->num_cus * 16));
        HIP_CHECK(launch_editops(a->width, ea, lev, nblocks, s));
        HIP_CHECK(hipMemcpyAsync(score.data(), ea.score, (size_t)items * 4, hipMemcpyDeviceToHost, s));
        if (row_words) HIP_CHECK(hipMemcpyAsync(rows.data(), ea.rows, row_words * 8, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    std::vector<uint64_t> dist(n);
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const EdPair& p = pairs[i];
        if (p.item < 0)
            dist[i] = (uint64_t)p.la + p.lb;
        else if (lev)
            dist[i] = score[p.item];
        else
            dist[i] = (uint64_t)p.la + p.lb - 2ull * score[p.item];
        offsets[i + 1] = offsets[i] + dist[i];
    }
    if (!ops) return;
    if (capacity < offsets[n])
        throw ApiError(RFA_ERR_INVALID_ARGUMENT,
                       "ops holds " + std::to_string(capacity) + " operations, " + std::to_string(offsets[n]) + " needed");
    for (uint64_t i = 0; i < n; ++i)
        if (dist[i]) walk(pairs[i], lev, rows.data(), dist[i], a, b, ops + offsets[i]);
}

} // namespace

rfa_status rfa_editops(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, int metric, uint64_t* offsets,
                       rfa_editop* ops, uint64_t capacity)
{
    if (!ctx || !a || !b || !offsets) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, a, b or offsets is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        begin_call(ctx);
        editops_impl(ctx, a, b, metric, offsets, ops, capacity);
    });
}
