// rfa_pralign.cpp -- rfa_partial_ratio_alignment: fuzz::partial_ratio_alignment for
// batches of pairs (fuzz_impl.hpp:196-230).  The host settles the pairs the reference
// answers before searching (a cutoff above 100, an empty string), groups the rest by
// the needle's word count, sizes each class's per-wave LDS and gives the pairs whose
// workspace exceeds it a slice of scratch; the device searches (rfa_pralign.hip) and
// writes the results in the caller's layout, copied back in one piece.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_internal.hpp"

using namespace rfa;

namespace {

static_assert(sizeof(rfa_score_alignment) == 40, "rfa_score_alignment is five 8-byte fields");

void pralign_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, double cutoff,
                  rfa_score_alignment* out)
{
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    const uint64_t n = a->count;
    if (n == 0) return;
    if (n > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 pairs");
    // classes by needle words: 1, 2, 3-4, 5-8, 9-16
    constexpr int NC = 5;
    std::vector<uint8_t> cls(n);
    std::vector<uint32_t> wsb(n); // workspace bytes with symbol tables (LDS candidates)
    uint64_t cnt[NC] = {0, 0, 0, 0, 0};
    uint32_t lds[NC] = {0, 0, 0, 0, 0};
    std::vector<uint64_t> trivial;
    std::vector<uint32_t> big;
    uint64_t big_ws = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t la = a->len(i), lb = b->len(i);
        const uint64_t l1 = std::min(la, lb), l2 = std::max(la, lb);
        if (cutoff > 100 || l1 == 0) { // settled here (the swap for a longer first string keeps them)
            cls[i] = 0xff;
            trivial.push_back(i);
            continue;
        }
        if (l1 > PBIG_MAX)
            throw ApiError(RFA_ERR_UNSUPPORTED, "partial_ratio_alignment of two strings longer than 4096 code units");
        if (l1 > PRALIGN_MAX) { // one wave per pair, k_pbig
            cls[i] = 0xfe;
            big.push_back((uint32_t)i);
            big_ws = std::max<uint64_t>(big_ws, pbig_ws_bytes((int)a->width, l1, l2));
            continue;
        }
        const uint64_t w = (l1 + 63) / 64;
        const int c = w <= 1 ? 0 : w <= 2 ? 1 : w <= 4 ? 2 : w <= 8 ? 3 : 4;
        cls[i] = (uint8_t)c;
        ++cnt[c];
        const uint64_t ws = pralign_ws_bytes(l1, l2, pralign_symbol_tables((int)a->width, l1));
        wsb[i] = (uint32_t)std::min<uint64_t>(ws, 0xffffffffull);
        // wider needles never fit the LDS cap
        if (c <= 2 && ws <= PRALIGN_LDS_CAP) lds[c] = std::max(lds[c], (uint32_t)ws);
    }
    const uint64_t nitems = n - trivial.size() - big.size();
    hipStream_t s = ctx->stream;
    if (nitems || !big.empty()) {
        // items (class-major) and scratch offsets in one pinned upload; the device takes LDS
        // exactly when the workspace fits the class's size (the same test)
        uint64_t first[NC + 1] = {0};
        for (int c = 0; c < NC; ++c) first[c + 1] = first[c] + cnt[c];
        const size_t o_offs = (nitems * 4 + 7) / 8 * 8, up = o_offs + nitems * 8;
        int slot = 0;
        uint8_t* h = (uint8_t*)ctx->staging.acquire(up, &slot);
        uint32_t* hitems = (uint32_t*)h;
        uint64_t* hoffs = (uint64_t*)(h + o_offs);
        uint64_t pos[NC];
        for (int c = 0; c < NC; ++c) pos[c] = first[c];
        uint64_t total = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const int c = cls[i];
            if (c >= 0xfe) continue;
            const uint64_t k = pos[c]++;
            hitems[k] = (uint32_t)i;
            hoffs[k] = total;
            if (wsb[i] > lds[c]) {
                const uint64_t la = a->len(i), lb = b->len(i);
                total += pralign_ws_bytes(std::min(la, lb), std::max(la, lb), false);
            }
        }
        const size_t o_big = (up + 255) / 256 * 256, o_out = o_big + (big.size() * 4 + 255) / 256 * 256;
        uint8_t* d = (uint8_t*)ctx->ed_items.get(o_out + n * 40);
        ctx->staging.copy(slot, d, up, s);
        PrAlignArgs pa{};
        pa.achars = a->d_chars;
        pa.aoff = a->d_off;
        pa.bchars = b->d_chars;
        pa.boff = b->d_off;
        pa.scratch = (uint8_t*)ctx->ed_rows.get(std::max<uint64_t>(total, 8));
        pa.cutoff = cutoff;
        pa.out = (uint64_t*)(d + o_out);
        const uint32_t words[NC] = {1, 2, 4, 8, 16};
        for (int c = 0; c < NC; ++c) {
            pa.items = (const uint32_t*)d + first[c];
            pa.scratch_off = (const uint64_t*)(d + o_offs) + first[c];
            pa.nitems = (uint32_t)cnt[c];
            pa.lds_bytes = lds[c];
            HIP_CHECK(launch_pralign(a->width, words[c], pa, s));
        }
        if (!big.empty()) {
            HIP_CHECK(hipMemcpyAsync(d + o_big, big.data(), big.size() * 4, hipMemcpyHostToDevice, s));
            PbigArgs ba{};
            ba.achars = a->d_chars;
            ba.aoff = a->d_off;
            ba.bchars = b->d_chars;
            ba.boff = b->d_off;
            ba.items = (const uint32_t*)(d + o_big);
            ba.nitems = big.size();
            ba.cutoff = cutoff;
            ba.al_out = pa.out;
            ba.ws_stride = (big_ws + 255) / 256 * 256;
            const uint32_t nb = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(std::min<uint64_t>(big.size(), (uint64_t)ctx->num_cus * 4), (1ull << 30) / ba.ws_stride));
            ba.ws = (uint8_t*)ctx->pb_ws.get(nb * ba.ws_stride);
            HIP_CHECK(launch_pbig(a->width, ba, nb, s));
            HIP_CHECK(hipStreamSynchronize(s)); // big is a host vector
        }
        HIP_CHECK(hipMemcpyAsync(out, pa.out, n * 40, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    for (uint64_t i : trivial) {
        const uint64_t la = a->len(i), lb = b->len(i);
        const uint64_t l1 = std::min(la, lb);
        if (cutoff > 100)
            out[i] = rfa_score_alignment{0.0, 0, l1, 0, l1};
        else
            out[i] = rfa_score_alignment{la == lb ? 100.0 : 0.0, 0, 0, 0, 0};
    }
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
