// rfa_editops.cpp -- rfa_editops: levenshtein_editops and lcs_seq_editops /
// indel_editops for batches of pairs (Levenshtein_impl.hpp:959-1015 and 1229-1257,
// LCSseq_impl.hpp:440-534).  The device strips the common affixes, records the bit
// rows and walks them back (rfa_editops.hip; Levenshtein pairs past 64 units through
// the block kernels, their rows in per-wave scratch); the host turns the per-pair
// counts into offsets and moves the operations.  Levenshtein pairs the reference splits
// by Hirschberg (levenshtein_align_hirschberg, Levenshtein_impl.hpp:1185-1212) are split
// level by level: the host strips each piece's affixes and sorts the pieces into leaves
// and splits, the device finds every split's position (one wave each), and all leaves
// are recorded and walked in one launch at the end.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_internal.hpp"

using namespace rfa;

namespace {

static_assert(sizeof(EdOp) == sizeof(rfa_editop), "EdOp mirrors rfa_editop");

// One piece of a split pair: s1[a0, a0 + la) against s2[b0, b0 + lb), its operations at
// opos, its distance at most max.
struct HbNode {
    uint32_t pair;
    uint64_t a0, la, b0, lb, src, dst, opos, max, dist;
};

// levenshtein_align_hirschberg (Levenshtein_impl.hpp:1185-1212) over every pair the device
// listed: each piece is stripped of its common affixes; a piece whose band matrix stays under
// 1 MiB (or whose s1 has under 65 units, or s2 under 10) becomes a leaf, the others are split
// at s2's middle and the s1 position the device finds (find_hirschberg_pos: the first
// smallest sum of the two rows).  Returns the leaves and raises `need` to their row area.
std::vector<EdLeaf> hirschberg_leaves(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b,
                                      const EditopsArgs& ea, const uint64_t* offsets, uint64_t nsplit,
                                      uint64_t& need)
{
    const uint64_t n = a->count;
    const auto host_copy = [](const rfa_strings* x) {
        return x->h_chars.size() >= (x->h_off[x->count] - x->h_off[0]) * (uint64_t)x->width;
    };
    if (!host_copy(a) || !host_copy(b))
        throw ApiError(RFA_ERR_UNSUPPORTED, "Levenshtein editops of pairs split by Hirschberg (|s1| > 64, |s2| >= 10 "
                                            "and |s1| * |s2| >= 2^22 without the common affixes) need string sets "
                                            "that hold a host copy, not views");
    hipStream_t s = ctx->stream;
    std::vector<uint32_t> list(nsplit);
    HIP_CHECK(hipMemcpyAsync(list.data(), ea.hlist, nsplit * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::sort(list.begin(), list.end());
    std::vector<HbNode> level;
    for (uint32_t p : list) {
        if (p >= n) throw ApiError(RFA_ERR_HIP, "Hirschberg pair index out of range");
        const uint64_t la = a->len(p), lb = b->len(p);
        level.push_back({p, 0, la, 0, lb, 0, 0, offsets[p], std::max(la, lb), offsets[p + 1] - offsets[p]});
    }
    std::vector<EdLeaf> leaves;
    std::vector<EdSplit> splits;
    std::vector<HbNode> pending;
    while (!level.empty()) {
        splits.clear();
        pending.clear();
        for (HbNode nd : level) {
            const uint64_t ab = a->h_off[nd.pair] - a->h_off[0] + nd.a0, bb = b->h_off[nd.pair] - b->h_off[0] + nd.b0;
            uint64_t pre = 0;
            while (pre < nd.la && pre < nd.lb && a->unit(ab + pre) == b->unit(bb + pre)) ++pre;
            uint64_t suf = 0;
            while (suf < nd.la - pre && suf < nd.lb - pre &&
                   a->unit(ab + nd.la - 1 - suf) == b->unit(bb + nd.lb - 1 - suf))
                ++suf;
            nd.a0 += pre;
            nd.b0 += pre;
            nd.src += pre;
            nd.dst += pre;
            nd.la -= pre + suf;
            nd.lb -= pre + suf;
            nd.max = std::min(nd.max, std::max(nd.la, nd.lb));
            const uint64_t full_band = std::min(nd.la, 2 * nd.max + 1);
            if ((2 * full_band * nd.lb) / 8 < 1024 * 1024 || nd.la < 65 || nd.lb < 10) {
                EdLeaf L{};
                L.pair = nd.pair;
                L.la = (uint32_t)nd.la;
                L.lb = (uint32_t)nd.lb;
                L.a0 = nd.a0;
                L.b0 = nd.b0;
                L.src = nd.src;
                L.dst = nd.dst;
                L.opos = nd.opos;
                L.dist = (uint32_t)nd.dist;
                if (nd.la && nd.lb) need = std::max<uint64_t>(need, (2 * ((nd.la + 63) / 64) + 1) * nd.lb);
                leaves.push_back(L);
                continue;
            }
            if (nd.la > EDITOPS_LEV_BLOCK_MAX || nd.lb > 0xffffffffull)
                throw ApiError(RFA_ERR_HIP, "Hirschberg piece past the block limit");
            EdSplit sp{};
            sp.pair = nd.pair;
            sp.la = (uint32_t)nd.la;
            sp.lb = (uint32_t)nd.lb;
            sp.a0 = nd.a0;
            sp.b0 = nd.b0;
            splits.push_back(sp);
            pending.push_back(nd);
        }
        level.clear();
        if (splits.empty()) break;
        EdSplit* d_splits = (EdSplit*)ctx->ed_hnodes.get(splits.size() * sizeof(EdSplit));
        const uint64_t tmp_stride = 2 * (EDITOPS_LEV_BLOCK_MAX + 1) + 62; // two rows of up to 4097 values
        const uint32_t nb = (uint32_t)std::min<uint64_t>(splits.size(), (uint64_t)ctx->num_cus * 8);
        uint32_t* tmp = (uint32_t*)ctx->ed_htmp.get(nb * tmp_stride * 4);
        HIP_CHECK(hipMemcpyAsync(d_splits, splits.data(), splits.size() * sizeof(EdSplit), hipMemcpyHostToDevice, s));
        HIP_CHECK(launch_editops_split(a->width, ea, d_splits, splits.size(), tmp, tmp_stride, nb, s));
        HIP_CHECK(hipMemcpyAsync(splits.data(), d_splits, splits.size() * sizeof(EdSplit), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (size_t k = 0; k < splits.size(); ++k) {
            const EdSplit& sp = splits[k];
            const HbNode& nd = pending[k];
            if (sp.mid > nd.la || (uint64_t)sp.left + sp.right != nd.dist)
                throw ApiError(RFA_ERR_HIP, "Hirschberg split out of range");
            const uint64_t ls = nd.lb / 2;
            level.push_back({nd.pair, nd.a0, sp.mid, nd.b0, ls, nd.src, nd.dst, nd.opos, sp.left, sp.left});
            level.push_back({nd.pair, nd.a0 + sp.mid, nd.la - sp.mid, nd.b0 + ls, nd.lb - ls, nd.src + sp.mid,
                             nd.dst + ls, nd.opos + sp.left, sp.right, sp.right});
        }
    }
    return leaves;
}

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
                 o_hcount = o_need + 8, o_hlist = o_hcount + 8,
                 o_temp = (o_hlist + (lev ? n * 4 : 0) + 255) / 256 * 256;
    uint8_t* items = (uint8_t*)ctx->ed_items.get(o_temp + scan_bytes);
    ea.info = (EdInfo*)(items + o_info);
    ea.counts = (uint64_t*)(items + o_counts);
    uint64_t* d_offsets = (uint64_t*)(items + o_offsets);
    ea.offsets = d_offsets;
    ea.err = (uint32_t*)(items + o_err);
    ea.need = (uint64_t*)(items + o_need);
    ea.hcount = (uint64_t*)(items + o_hcount);
    ea.hlist = (uint32_t*)(items + o_hlist);
    ea.n = n;
    if (!lev && !ham) {
        ea.row_base = (const uint64_t*)(items + o_base);
        HIP_CHECK(hipMemcpyAsync((void*)ea.row_base, row_base.data(), n * 8, hipMemcpyHostToDevice, s));
    }
    HIP_CHECK(hipMemsetAsync(ea.counts + n, 0, 8, s));
    HIP_CHECK(hipMemsetAsync(ea.err, 0, 24, s)); // the error word, the row area, the split count
    const uint32_t nblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 3) / 4, (uint64_t)ctx->num_cus * 16));
    // Levenshtein pairs past 64 units: the waves scan for them round robin, up to a wave per pair
    const uint32_t lblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 3) / 4, (uint64_t)ctx->num_cus * 8));
    if (ham)
        HIP_CHECK(launch_editops_hamming(a->width, ea, false, nblocks, s));
    else
        HIP_CHECK(launch_editops(a->width, ea, lev, false, nblocks, s));
    if (lev) HIP_CHECK(launch_editops_lev_block(a->width, ea, false, lblocks, s));
    HIP_CHECK(editops_offsets(ea.counts, d_offsets, n, items + o_temp, scan_bytes, s));
    uint64_t errneed[3] = {0, 0, 0};
    HIP_CHECK(hipMemcpyAsync(errneed, ea.err, 24, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(offsets, d_offsets, (n + 1) * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if ((uint32_t)errneed[0])
        throw ApiError(RFA_ERR_UNSUPPORTED,
                       lev ? "Levenshtein editops of a pair whose first string has more than 4096 code units "
                             "without the common prefix and suffix"
                           : "LCS editops of a pair whose first string has more than 4096 code units without "
                             "the common prefix and suffix");
    uint64_t need = errneed[1];
    const uint64_t nsplit = errneed[2];
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
    std::vector<EdLeaf> leaves;
    if (lev && nsplit) leaves = hirschberg_leaves(ctx, a, b, ea, offsets, nsplit, need);
    if (lev && need) {
        // the block pairs' and the leaves' rows: one piece's area per wave, at most 1 GiB in all
        ea.long_stride = (need + 31) / 32 * 32;
        const uint64_t per_block = 4 * ea.long_stride * 8;
        const uint32_t wb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(lblocks, (1ull << 30) / per_block));
        ea.long_rows = (uint64_t*)ctx->ed_long.get(wb * per_block);
        HIP_CHECK(launch_editops_lev_block(a->width, ea, true, wb, s));
        if (!leaves.empty()) {
            EdLeaf* d_leaves = (EdLeaf*)ctx->ed_hnodes.get(leaves.size() * sizeof(EdLeaf));
            HIP_CHECK(hipMemcpyAsync(d_leaves, leaves.data(), leaves.size() * sizeof(EdLeaf), hipMemcpyHostToDevice, s));
            ea.leaves = d_leaves;
            ea.nleaves = leaves.size();
            const uint32_t lb2 = (uint32_t)std::min<uint64_t>(wb, (leaves.size() + 3) / 4);
            HIP_CHECK(launch_editops_lev_block(a->width, ea, true, lb2, s));
            ea.leaves = nullptr;
            ea.nleaves = 0;
        }
    }
    HIP_CHECK(hipMemcpyAsync(ops, ea.ops, total * sizeof(EdOp), hipMemcpyDeviceToHost, s));
    uint32_t werr = 0;
    if (lev) HIP_CHECK(hipMemcpyAsync(&werr, ea.err, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (werr) throw ApiError(RFA_ERR_HIP, "Levenshtein editops: a block walk's distance differs from its count");
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
