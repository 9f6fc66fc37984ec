// rfa_pralign.cpp -- rfa_partial_ratio_alignment: fuzz::partial_ratio_alignment for
// batches of pairs (fuzz_impl.hpp:196-230).  The host settles the pairs the reference
// answers before searching (a cutoff above 100, an empty string), groups the rest by
// the needle's word count and gives each pair a slice of scratch; the device searches
// (rfa_pralign.hip).
#include <algorithm>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_internal.hpp"

using namespace rfa;

namespace {

void pralign_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, double cutoff,
                  rfa_score_alignment* out)
{
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    const uint64_t n = a->count;
    if (n == 0) return;
    if (n > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 pairs");
    // items by class (needle words 1, 2-4, 5-16), each with its scratch slice
    std::vector<uint32_t> items[3];
    std::vector<uint64_t> offs[3];
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t la = a->len(i), lb = b->len(i);
        const uint64_t l1 = std::min(la, lb), l2 = std::max(la, lb);
        rfa_score_alignment& r = out[i];
        // partial_ratio_alignment: the swap for a longer first string leaves these as they are
        if (cutoff > 100) {
            r = rfa_score_alignment{0.0, 0, l1, 0, l1};
            continue;
        }
        if (l1 == 0) {
            r = rfa_score_alignment{la == lb ? 100.0 : 0.0, 0, 0, 0, 0};
            continue;
        }
        if (l1 > PRALIGN_MAX)
            throw ApiError(RFA_ERR_UNSUPPORTED, "partial_ratio_alignment of two strings longer than 1024 code units");
        const uint64_t w = (l1 + 63) / 64;
        const int c = w <= 1 ? 0 : (w <= 4 ? 1 : 2);
        items[c].push_back((uint32_t)i);
        offs[c].push_back(total);
        total += pralign_scratch_bytes(l1, l2);
    }
    uint64_t nitems = items[0].size() + items[1].size() + items[2].size();
    if (nitems == 0) return;
    hipStream_t s = ctx->stream;
    // device: items + offsets of every class, scores and windows of every pair, scratch
    std::vector<uint32_t> all_items;
    std::vector<uint64_t> all_offs;
    for (int c = 0; c < 3; ++c) {
        all_items.insert(all_items.end(), items[c].begin(), items[c].end());
        all_offs.insert(all_offs.end(), offs[c].begin(), offs[c].end());
    }
    const size_t o_items = 0, o_offs = (nitems * 4 + 7) / 8 * 8, o_score = o_offs + nitems * 8,
                 o_al = o_score + n * 8, o_end = o_al + n * 32;
    uint8_t* d = (uint8_t*)ctx->ed_items.get(o_end);
    HIP_CHECK(hipMemcpyAsync(d + o_items, all_items.data(), nitems * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d + o_offs, all_offs.data(), nitems * 8, hipMemcpyHostToDevice, s));
    PrAlignArgs pa{};
    pa.achars = a->d_chars;
    pa.aoff = a->d_off;
    pa.bchars = b->d_chars;
    pa.boff = b->d_off;
    pa.scratch = (uint8_t*)ctx->ed_rows.get(std::max<uint64_t>(total, 8));
    pa.cutoff = cutoff;
    pa.score = (double*)(d + o_score);
    pa.al = (uint64_t*)(d + o_al);
    uint64_t first = 0;
    const uint32_t words[3] = {1, 4, 16};
    for (int c = 0; c < 3; ++c) {
        pa.items = (const uint32_t*)(d + o_items) + first;
        pa.scratch_off = (const uint64_t*)(d + o_offs) + first;
        pa.nitems = (uint32_t)items[c].size();
        HIP_CHECK(launch_pralign(a->width, words[c], pa, s));
        first += items[c].size();
    }
    std::vector<double> score(n);
    std::vector<uint64_t> al(4 * n);
    HIP_CHECK(hipMemcpyAsync(score.data(), pa.score, n * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(al.data(), pa.al, n * 32, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int c = 0; c < 3; ++c)
        for (uint32_t i : items[c])
            out[i] = rfa_score_alignment{score[i], al[4 * i], al[4 * i + 1], al[4 * i + 2], al[4 * i + 3]};
}

} // namespace

rfa_status rfa_partial_ratio_alignment(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b,
                                       double score_cutoff, rfa_score_alignment* out)
{
    if (!ctx || !a || !b || !out) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, a, b or out is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        begin_call(ctx);
        pralign_impl(ctx, a, b, score_cutoff, out);
    });
}
