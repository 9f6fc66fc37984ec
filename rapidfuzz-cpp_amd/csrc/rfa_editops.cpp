// rfa_editops.cpp -- rfa_editops: levenshtein_editops and lcs_seq_editops /
// indel_editops for batches of pairs (Levenshtein_impl.hpp:959-1015 and 1229-1257,
// LCSseq_impl.hpp:440-534).  The device strips the common affixes, records the bit
// rows and walks them back (rfa_editops.hip; Levenshtein pairs past 64 units through
// the block kernels, their rows in per-wave scratch); the host turns the per-pair
// counts into offsets and moves the operations.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_internal.hpp"

using namespace rfa;

namespace {

static_assert(sizeof(EdOp) == sizeof(rfa_editop), "EdOp mirrors rfa_editop");

void editops_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, int metric, uint64_t* offsets,
                  rfa_editop* ops, uint64_t capacity)
{
    if (metric != RFA_LEVENSHTEIN && metric != RFA_INDEL && metric != RFA_LCSSEQ && metric != RFA_HAMMING)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "editops are for Levenshtein, Indel, LCSseq and Hamming");
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    const uint64_t n = a->count;
    offsets[0] = 0;
    if (n == 0) return;
    const bool lev = metric == RFA_LEVENSHTEIN;
    const bool ham = metric == RFA_HAMMING;
    hipStream_t s = ctx->stream;
    EditopsArgs ea{};
    ea.achars = a->d_chars;
    ea.bchars = b->d_chars;
    ea.aoff = a->d_off;
    ea.boff = b->d_off;
    ea.b_base = b->h_off[0];
    // rows: Levenshtein two words per text unit (laid out by b's offsets); LCS
    // ceil(|a[i]| / 64) words per text unit (an upper bound for the stripped strings)
    uint64_t row_words = 0;
    std::vector<uint64_t> row_base;
    if (ham) {
        // no rows
    }
    else if (lev) {
        row_words = 2ull * (b->h_off[n] - b->h_off[0]);
    }
    else {
        row_base.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            row_base[i] = row_words;
            const uint64_t la = a->len(i), lb = b->len(i);
            row_words += std::min<uint64_t>((la + 63) / 64, EDITOPS_LCS_MAX / 64) * lb;
        }
    }
    ea.rows = (uint64_t*)ctx->ed_rows.get(std::max<uint64_t>(8, row_words * 8));
    // items: info, counts (n + 1), offsets (n + 1), row bases (LCS), the error word, the
    // block pairs' largest row area, scan temp
    const size_t scan_bytes = editops_scan_bytes(n);
    const size_t o_info = 0, o_counts = n * sizeof(EdInfo), o_offsets = o_counts + (n + 1) * 8,
                 o_base = o_offsets + (n + 1) * 8, o_err = o_base + (lev || ham ? 0 : n * 8), o_need = o_err + 8,
                 o_temp = (o_need + 8 + 255) / 256 * 256;
    uint8_t* items = (uint8_t*)ctx->ed_items.get(o_temp + scan_bytes);
    ea.info = (EdInfo*)(items + o_info);
    ea.counts = (uint64_t*)(items + o_counts);
    uint64_t* d_offsets = (uint64_t*)(items + o_offsets);
    ea.offsets = d_offsets;
    ea.err = (uint32_t*)(items + o_err);
    ea.need = (uint64_t*)(items + o_need);
    ea.n = n;
    if (!lev && !ham) {
        ea.row_base = (const uint64_t*)(items + o_base);
        HIP_CHECK(hipMemcpyAsync((void*)ea.row_base, row_base.data(), n * 8, hipMemcpyHostToDevice, s));
    }
    HIP_CHECK(hipMemsetAsync(ea.counts + n, 0, 8, s));
    HIP_CHECK(hipMemsetAsync(ea.err, 0, 16, s)); // the error word and the row area
    const uint32_t nblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 3) / 4, (uint64_t)ctx->num_cus * 16));
    // Levenshtein pairs past 64 units: one wave per 64 pairs scans for them
    const uint32_t lblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->num_cus * 8));
    if (ham)
        HIP_CHECK(launch_editops_hamming(a->width, ea, false, nblocks, s));
    else
        HIP_CHECK(launch_editops(a->width, ea, lev, false, nblocks, s));
    if (lev) HIP_CHECK(launch_editops_lev_block(a->width, ea, false, lblocks, s));
    HIP_CHECK(editops_offsets(ea.counts, d_offsets, n, items + o_temp, scan_bytes, s));
    uint64_t errneed[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(errneed, ea.err, 16, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(offsets, d_offsets, (n + 1) * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if ((uint32_t)errneed[0])
        throw ApiError(RFA_ERR_UNSUPPORTED,
                       lev ? "Levenshtein editops of a pair whose first string, without the common prefix and "
                             "suffix, has more than 4096 code units, or whose matrix the reference splits "
                             "(Hirschberg: 64 < |s1| and |s1| * |s2| >= 2^22, |s2| >= 10)"
                           : "LCS editops of a pair whose first string has more than 4096 code units without "
                             "the common prefix and suffix");
    const uint64_t need = errneed[1];
    if (!ops) return;
    const uint64_t total = offsets[n];
    if (capacity < total)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT,
                       "ops holds " + std::to_string(capacity) + " operations, " + std::to_string(total) + " needed");
    if (total == 0) return;
    ea.ops = (EdOp*)ctx->ed_ops.get(total * sizeof(EdOp));
    if (ham)
        HIP_CHECK(launch_editops_hamming(a->width, ea, true, nblocks, s));
    else
        HIP_CHECK(launch_editops(a->width, ea, lev, true, 0, s));
    if (lev && need) {
        // the block pairs' rows: one pair's area per wave, at most 1 GiB in all
        ea.long_stride = (need + 31) / 32 * 32;
        const uint64_t per_block = 4 * ea.long_stride * 8;
        const uint32_t wb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(lblocks, (1ull << 30) / per_block));
        ea.long_rows = (uint64_t*)ctx->ed_long.get(wb * per_block);
        HIP_CHECK(launch_editops_lev_block(a->width, ea, true, wb, s));
    }
    HIP_CHECK(hipMemcpyAsync(ops, ea.ops, total * sizeof(EdOp), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
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
