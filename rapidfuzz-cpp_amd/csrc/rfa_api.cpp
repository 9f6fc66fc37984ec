// rfa_api.cpp -- the C ABI: contexts, resident string sets, scorers, and the
// host-side planning that splits a request into kernel launches.
//
// Planning (done once per scorer, like the reference's PM construction in
// the Cached* constructors):
//   * queries of <= 64 code units are sorted by length (longest first) and cut
//     into tiles of 64; a tile whose queries all fit in 32 units runs with
//     32-bit words.  Each tile's distinct-symbol count sizes its LDS PM table.
//   * longer queries (and tiles whose alphabet does not fit in LDS) are scored
//     by the block kernel, one wave per pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "rfa_host.hpp"
#include "rfa_semantics.hpp"

namespace rfa {

thread_local std::string g_err;

// Called at the start of every scoring call: releases the host sources of the
// previous call's uploads once the device has consumed them.
void begin_call(rfa_context* ctx)
{
    if (!ctx->keep.items.empty()) {
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        HIP_CHECK(hipStreamSynchronize(ctx->stream2));
        ctx->keep.items.clear();
    }
}

int algo_for(int metric, const rfa_weights& w)
{
    switch (metric) {
    case RFA_LEVENSHTEIN:
        if (w.insert_cost == w.delete_cost) {
            if (w.insert_cost == 0 || w.insert_cost == w.replace_cost) return ALGO_LEV;
            if (w.replace_cost >= w.insert_cost + w.delete_cost) return ALGO_LCS;
        }
        return ALGO_WLEV;
    case RFA_INDEL:
    case RFA_LCSSEQ:
    case RFA_RATIO: return ALGO_LCS;
    case RFA_OSA: return ALGO_OSA;
    case RFA_HAMMING:
    case RFA_PREFIX:
    case RFA_POSTFIX: return ALGO_POSIT;
    case RFA_PARTIAL_RATIO:
    case RFA_PARTIAL_TOKEN_SORT_RATIO: return ALGO_PARTIAL;
    case RFA_TOKEN_SORT_RATIO:
    case RFA_QRATIO: return ALGO_LCS;
    case RFA_TOKEN_SET_RATIO:
    case RFA_TOKEN_RATIO:
    case RFA_PARTIAL_TOKEN_SET_RATIO:
    case RFA_PARTIAL_TOKEN_RATIO: return ALGO_TSET;
    case RFA_JARO:
    case RFA_JARO_WINKLER: return ALGO_JARO;
    case RFA_DAMERAU_LEVENSHTEIN: return ALGO_DL;
    case RFA_WRATIO: return ALGO_LCS;
    }
    return -1;
}

bool multi_levenshtein_weights(const rfa_weights& w)
{
    return w.insert_cost == 1 && w.delete_cost == 1 && w.replace_cost <= 2;
}

bool is_int_kind(int metric, int kind)
{
    return (metric == RFA_LEVENSHTEIN || metric == RFA_INDEL || metric == RFA_LCSSEQ || metric == RFA_OSA ||
            metric == RFA_HAMMING || metric == RFA_PREFIX || metric == RFA_POSTFIX ||
            metric == RFA_DAMERAU_LEVENSHTEIN) &&
           (kind == RFA_DISTANCE || kind == RFA_SIMILARITY);
}

// The distinct-symbol count above which a query tile leaves the LDS path.
constexpr uint32_t TILE_MAX_ROWS = 255;

void plan_scorer(rfa_scorer* sc)
{
    rfa_strings* q = sc->q;
    std::vector<uint32_t> shortq;
    for (uint64_t i = 0; i < q->count; ++i) {
        if (q->len(i) <= 64)
            shortq.push_back((uint32_t)i);
        else
            sc->block_queries.push_back((uint32_t)i);
    }
    std::stable_sort(shortq.begin(), shortq.end(), [&](uint32_t a, uint32_t b) { return q->len(a) > q->len(b); });
    // tiles; a tile whose alphabet is too large for LDS moves to the block path
    std::vector<uint32_t> kept;
    std::vector<uint32_t> rows_of_tile;
    for (size_t t0 = 0; t0 < shortq.size(); t0 += 64) {
        size_t t1 = std::min(shortq.size(), t0 + 64);
        std::unordered_set<uint32_t> syms;
        for (size_t k = t0; k < t1; ++k) {
            uint32_t qi = shortq[k];
            for (uint64_t p = q->h_off[qi]; p < q->h_off[qi + 1]; ++p) syms.insert(q->unit(p));
        }
        if (syms.size() + 1 > TILE_MAX_ROWS) {
            for (size_t k = t0; k < t1; ++k) sc->block_queries.push_back(shortq[k]);
            continue;
        }
        for (size_t k = t0; k < t1; ++k) kept.push_back(shortq[k]);
        rows_of_tile.push_back((uint32_t)syms.size() + 1);
    }
    // kept stays sorted by length (whole tiles were removed), so tiles are re-cut identically
    // only if no tile was removed before a kept one; re-cut to keep tiles full.
    std::vector<uint32_t> order = kept;
    std::vector<uint32_t> rows;
    for (size_t t0 = 0; t0 < order.size(); t0 += 64) {
        size_t t1 = std::min(order.size(), t0 + 64);
        std::unordered_set<uint32_t> syms;
        for (size_t k = t0; k < t1; ++k) {
            uint32_t qi = order[k];
            for (uint64_t p = q->h_off[qi]; p < q->h_off[qi + 1]; ++p) syms.insert(q->unit(p));
        }
        rows.push_back((uint32_t)syms.size() + 1);
    }
    sc->tile_order = order;
    uint32_t ntiles = (uint32_t)rows.size();
    uint32_t split = ntiles; // first tile with all queries <= 32 units
    for (uint32_t t = 0; t < ntiles; ++t) {
        if (q->len(order[(size_t)t * 64]) <= 32) {
            split = t;
            break;
        }
    }
    uint32_t max_rows_all = 2;
    for (int c = 0; c < 2; ++c) {
        TileClass tc;
        tc.first_tile = c == 0 ? 0 : split;
        tc.ntiles = c == 0 ? split : ntiles - split;
        tc.word_bits = c == 0 ? 64 : 32;
        uint32_t mr = 2;
        for (uint32_t t = tc.first_tile; t < tc.first_tile + tc.ntiles; ++t) mr = std::max(mr, rows[t]);
        tc.rows_cap = mr;
        max_rows_all = std::max(max_rows_all, mr);
        if (tc.ntiles) sc->classes.push_back(tc);
    }
    sc->hash_bits = std::max<uint32_t>(6, log2_ceil(2ull * max_rows_all));
    for (uint32_t qi : sc->block_queries) sc->block_max_len = std::max(sc->block_max_len, q->len(qi));
    sc->block_min_len = sc->block_max_len;
    for (uint32_t qi : sc->block_queries) sc->block_min_len = std::min(sc->block_min_len, q->len(qi));
    if (q->width == 1) { // 8-bit block patterns are sized by their distinct symbols
        uint32_t rows = 0;
        sc->q_rows.assign(q->count, 0);
        for (uint32_t qi : sc->block_queries) {
            bool seen[256] = {};
            uint32_t d = 0;
            for (uint64_t pos = q->h_off[qi]; pos < q->h_off[qi + 1]; ++pos)
                if (!seen[q->h_chars[pos]]) seen[q->h_chars[pos]] = true, ++d;
            rows = std::max(rows, d + 1);
            sc->q_rows[qi] = (uint16_t)(d + 1);
        }
        sc->block_u8_rows = rows;
    }
    // longest first: the block queue is then longest-processing-time-first
    std::stable_sort(sc->block_queries.begin(), sc->block_queries.end(),
                     [&](uint32_t a, uint32_t b) { return q->len(a) > q->len(b); });
    if (!order.empty()) {
        HIP_CHECK(hipMalloc(&sc->d_tile_order, order.size() * sizeof(uint32_t)));
        HIP_CHECK(hipMemcpy(sc->d_tile_order, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    if (!sc->block_queries.empty()) {
        HIP_CHECK(hipMalloc(&sc->d_block_queries, sc->block_queries.size() * sizeof(uint32_t)));
        HIP_CHECK(hipMemcpy(sc->d_block_queries, sc->block_queries.data(), sc->block_queries.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice));
    }
}

rfa_strings* make_strings(rfa_context* ctx, const void* chars, int width, const uint64_t* off, uint64_t count)
{
    if (width != 1 && width != 2 && width != 4) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "char_width must be 1, 2 or 4");
    if (!off && count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "offsets is NULL");
    static std::atomic<uint64_t> next_id{1};
    auto* s = new rfa_strings();
    s->ctx = ctx;
    s->width = width;
    s->count = count;
    s->id = next_id.fetch_add(1);
    s->h_off.assign(off, off + count + 1);
    if (count == 0) s->h_off.assign(1, 0);
    for (uint64_t i = 0; i < count; ++i) {
        if (s->h_off[i + 1] < s->h_off[i]) {
            delete s;
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
        }
        s->max_len = std::max(s->max_len, s->h_off[i + 1] - s->h_off[i]);
    }
    uint64_t nunits = s->h_off.back();
    if (nunits && !chars) {
        delete s;
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "chars is NULL");
    }
    if (s->max_len > 0xffffffffull) {
        // the kernels count a string's units in 32 bits
        delete s;
        throw ApiError(RFA_ERR_UNSUPPORTED, "strings longer than 2^32 - 1 code units");
    }
    s->h_chars.resize(nunits * width);
    if (nunits) std::memcpy(s->h_chars.data(), chars, nunits * width);
    // 64 bytes of tail padding: kernels load code units in short vectors and may read
    // a few units past the last string
    size_t cb = nunits * width + 64;
    if (hipMalloc(&s->d_chars, cb) != hipSuccess || hipMalloc((void**)&s->d_off, s->h_off.size() * 8) != hipSuccess) {
        (void)hipGetLastError();
        if (s->d_chars) (void)hipFree(s->d_chars);
        delete s;
        throw ApiError(RFA_ERR_OUT_OF_MEMORY, "device allocation for strings failed");
    }
    try {
        if (nunits) HIP_CHECK(hipMemcpy(s->d_chars, s->h_chars.data(), nunits * width, hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(s->d_off, s->h_off.data(), s->h_off.size() * 8, hipMemcpyHostToDevice));
    }
    catch (...) {
        free_strings(s);
        throw;
    }
    return s;
}

void free_strings(rfa_strings* s)
{
    if (!s) return;
    if (!s->view) {
        if (s->d_chars) (void)hipFree(s->d_chars);
        if (s->d_off) (void)hipFree(s->d_off);
    }
    delete s;
}

// token-sorted copy of a set; with `table`, deduplicated, and the token table filled in
static rfa_strings* token_transform(rfa_context* ctx, const rfa_strings* in, bool signed8, TokenTable* table)
{
    hipStream_t s = ctx->stream;
    const uint64_t n = in->count, w = (uint64_t)in->width;
    const uint64_t units = in->h_off.back();
    const size_t scan = tokens_scan_bytes(n);
    // one allocation: out_len, ntok, out_off, tok_off (n + 1 each), tok_start, tok_len, dup, out units,
    // scan temp; the token table (set_off, hash, start, len) in its own allocation
    const size_t cnt = (size_t)(n + 1);
    const size_t tok = (size_t)std::max<uint64_t>(1, units);
    size_t bytes = 4 * cnt * 8 + 2 * tok * 4 + tok + tok * w + 64 + 256 + scan + 256;
    uint8_t* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw ApiError(RFA_ERR_OUT_OF_MEMORY, "device allocation for the token-sorted strings failed");
    }
    TokenArgs a{};
    a.chars = in->d_chars;
    a.off = in->d_off;
    a.count = n;
    a.out_len = (uint64_t*)base;
    a.ntok = a.out_len + cnt;
    a.out_off = a.ntok + cnt;
    a.tok_off = a.out_off + cnt;
    a.tok_start = (uint32_t*)(a.tok_off + cnt);
    a.tok_len = a.tok_start + tok;
    a.dup = (uint8_t*)(a.tok_len + tok);
    a.out = (void*)(((uintptr_t)(a.dup + tok) + 7) & ~(uintptr_t)7);
    a.scan_temp = (void*)(((uintptr_t)((uint8_t*)a.out + tok * w + 64) + 255) & ~(uintptr_t)255);
    a.scan_bytes = scan;
    a.signed8 = signed8 ? 1 : 0;
    a.num_cus = (uint32_t)ctx->num_cus;
    if (table) {
        // kept tokens: at most (units + 1) / 2
        const size_t kt = (size_t)(units + 1) / 2 + 1;
        if (hipMalloc(&table->d_blob, cnt * 8 + kt * 16) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(base);
            throw ApiError(RFA_ERR_OUT_OF_MEMORY, "device allocation for the token table failed");
        }
        table->d_off = (uint64_t*)table->d_blob;
        table->d_hash = table->d_off + cnt;
        table->d_start = (uint32_t*)(table->d_hash + kt);
        table->d_len = table->d_start + kt;
        a.dedup = 1;
        a.set_off = table->d_off;
        a.set_hash = table->d_hash;
        a.set_start = table->d_start;
        a.set_len = table->d_len;
    }
    std::vector<uint64_t> h_off(cnt);
    std::vector<uint8_t> h_chars;
    hipError_t e = hipMemsetAsync(a.out_len + n, 0, 8, s);
    if (e == hipSuccess) e = hipMemsetAsync(a.ntok + n, 0, 8, s);
    if (e == hipSuccess) e = launch_tokens(in->width, a, s);
    if (e == hipSuccess) e = hipMemcpyAsync(h_off.data(), a.out_off, cnt * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) {
        h_chars.resize(h_off[n] * w);
        if (!h_chars.empty()) e = hipMemcpy(h_chars.data(), a.out, h_chars.size(), hipMemcpyDeviceToHost);
    }
    (void)hipFree(base);
    if (e != hipSuccess) throw ApiError(RFA_ERR_HIP, std::string("token sort: ") + hipGetErrorString(e));
    return make_strings(ctx, h_chars.data(), in->width, h_off.data(), n);
}

rfa_strings* token_sorted(rfa_context* ctx, const rfa_strings* in, bool signed8)
{
    return token_transform(ctx, in, signed8, nullptr);
}

TokenTable::~TokenTable()
{
    free_strings(joined);
    if (d_blob) (void)hipFree(d_blob);
}

std::shared_ptr<TokenTable> token_table(rfa_context* ctx, const rfa_strings* in, bool signed8)
{
    auto t = std::make_shared<TokenTable>();
    t->joined = token_transform(ctx, in, signed8, t.get());
    return t;
}

const rfa_strings* scorer_choices(rfa_scorer* sc, const rfa_strings* ch)
{
    if (sc->token_set) {
        if (!sc->c_tokens || sc->c_tokens_id != ch->id || sc->c_tokens_signed8 != sc->signed8) {
            sc->c_tokens = token_table(sc->ctx, ch, sc->signed8);
            sc->c_tokens_id = ch->id;
            sc->c_tokens_signed8 = sc->signed8;
        }
        return sc->c_tokens->joined;
    }
    if (!sc->token_sort) return ch;
    if (!sc->tok_choices || sc->tok_choices_id != ch->id || sc->tok_choices_signed8 != sc->signed8) {
        rfa_strings* t = token_sorted(sc->ctx, ch, sc->signed8);
        free_strings(sc->tok_choices);
        sc->tok_choices = t;
        sc->tok_choices_id = ch->id;
        sc->tok_choices_signed8 = sc->signed8;
    }
    return sc->tok_choices;
}

// drops a scorer's plan (its query set and tile / block lists)
static void clear_plan(rfa_scorer* sc)
{
    free_strings(sc->q);
    sc->q = nullptr;
    if (sc->d_tile_order) (void)hipFree(sc->d_tile_order);
    if (sc->d_block_queries) (void)hipFree(sc->d_block_queries);
    sc->d_tile_order = nullptr;
    sc->d_block_queries = nullptr;
    sc->tile_order.clear();
    sc->classes.clear();
    sc->block_queries.clear();
    sc->block_max_len = sc->block_min_len = 0;
    sc->block_u8_rows = 0;
    sc->q_rows.clear();
    sc->hash_bits = 6;
    sc->few_plan.reset();
}

rfa_scorer* make_scorer(rfa_context* ctx, int metric, const rfa_strings* queries, const rfa_weights& w,
                        double prefix_weight)
{
    auto* sc = new rfa_scorer();
    sc->ctx = ctx;
    sc->api_metric = metric;
    sc->metric = metric;
    sc->w = w;
    sc->prefix_weight = prefix_weight;
    // token-sort scorers run ratio / partial_ratio on token-sorted copies, QRatio runs ratio
    if (metric == RFA_TOKEN_SORT_RATIO || metric == RFA_PARTIAL_TOKEN_SORT_RATIO) {
        sc->metric = metric == RFA_TOKEN_SORT_RATIO ? RFA_RATIO : RFA_PARTIAL_RATIO;
        sc->token_sort = true;
    }
    else if (metric == RFA_QRATIO) {
        sc->metric = RFA_RATIO;
        sc->empty_zero = true;
    }
    else if (metric == RFA_TOKEN_SET_RATIO || metric == RFA_TOKEN_RATIO || metric == RFA_PARTIAL_TOKEN_SET_RATIO ||
             metric == RFA_PARTIAL_TOKEN_RATIO) {
        sc->token_set = true;
    }
    else if (metric == RFA_WRATIO) {
        sc->metric = RFA_RATIO; // the kinds and semantics of a ratio scorer
        sc->wratio = true;
    }
    try {
        if (sc->wratio) {
            static const int subs[4] = {RFA_RATIO, RFA_TOKEN_RATIO, RFA_PARTIAL_RATIO, RFA_PARTIAL_TOKEN_RATIO};
            for (int i = 0; i < 4; ++i) sc->wr_sub[i] = make_scorer(ctx, subs[i], queries, w, prefix_weight);
            sc->q = make_strings(ctx, queries->h_chars.data(), queries->width, queries->h_off.data(), queries->count);
        }
        else if (sc->token_set) {
            sc->q_raw =
                make_strings(ctx, queries->h_chars.data(), queries->width, queries->h_off.data(), queries->count);
            sc->q_tokens = token_table(ctx, sc->q_raw, false);
            sc->q = sc->q_tokens->joined; // the plan runs over the deduplicated joins
            sc->q_tokens->joined = nullptr;
            if (metric == RFA_TOKEN_RATIO) sc->sort_sub = make_scorer(ctx, RFA_TOKEN_SORT_RATIO, queries, w, prefix_weight);
            if (metric == RFA_PARTIAL_TOKEN_RATIO)
                sc->sort_sub = make_scorer(ctx, RFA_PARTIAL_TOKEN_SORT_RATIO, queries, w, prefix_weight);
            if (metric == RFA_PARTIAL_TOKEN_SET_RATIO || metric == RFA_PARTIAL_TOKEN_RATIO)
                sc->pr_dedup = make_scorer(ctx, RFA_PARTIAL_RATIO, sc->q, w, prefix_weight);
        }
        else if (sc->token_sort) {
            sc->q_raw =
                make_strings(ctx, queries->h_chars.data(), queries->width, queries->h_off.data(), queries->count);
            sc->q = token_sorted(ctx, sc->q_raw, false);
        }
        else {
            sc->q = make_strings(ctx, queries->h_chars.data(), queries->width, queries->h_off.data(), queries->count);
        }
        plan_scorer(sc);
    }
    catch (...) {
        free_scorer(sc);
        throw;
    }
    return sc;
}

void free_scorer(rfa_scorer* sc)
{
    if (!sc) return;
    clear_plan(sc);
    free_strings(sc->q_raw);
    free_strings(sc->tok_choices);
    free_scorer(sc->sort_sub);
    free_scorer(sc->pr_dedup);
    free_scorer(sc->pr_dedup_dupq);
    for (rfa_scorer* sub : sc->wr_sub) free_scorer(sub);
    if (sc->d_dup_queries) (void)hipFree(sc->d_dup_queries);
    delete sc;
}

// RFA_SCORER_SIGNED_UNITS on a token-sort scorer: the queries are sorted again
void scorer_set_signed(rfa_scorer* sc, bool on)
{
    if (sc->signed8 == on) return;
    sc->signed8 = on;
    for (rfa_scorer* sub : sc->wr_sub)
        if (sub) scorer_set_signed(sub, on);
    if (sc->sort_sub) scorer_set_signed(sc->sort_sub, on);
    if (sc->token_set) {
        clear_plan(sc);
        sc->q_tokens = token_table(sc->ctx, sc->q_raw, on);
        sc->q = sc->q_tokens->joined;
        sc->q_tokens->joined = nullptr;
        plan_scorer(sc);
        free_scorer(sc->pr_dedup_dupq);
        sc->pr_dedup_dupq = nullptr;
        sc->dup_queries.clear();
        if (sc->d_dup_queries) (void)hipFree(sc->d_dup_queries);
        sc->d_dup_queries = nullptr;
        sc->dup_queries_ready = false;
        if (sc->pr_dedup) {
            free_scorer(sc->pr_dedup);
            sc->pr_dedup = nullptr;
            sc->pr_dedup = make_scorer(sc->ctx, RFA_PARTIAL_RATIO, sc->q, sc->w, sc->prefix_weight);
        }
        return;
    }
    if (!sc->token_sort) return;
    clear_plan(sc);
    sc->q = token_sorted(sc->ctx, sc->q_raw, on);
    plan_scorer(sc);
}

ScoreSpec make_spec_m(int metric, const rfa_weights& w, double pw, int kind, int semantics, uint64_t cu, double cd,
                      int algo)
{
    ScoreSpec s{};
    s.metric = metric;
    s.kind = kind;
    s.semantics = semantics;
    s.algo = algo;
    s.cutoff_u = cu;
    s.cutoff_d = cd;
    s.ins = w.insert_cost;
    s.del = w.delete_cost;
    s.rep = w.replace_cost;
    s.prefix_weight = pw;
    s.band_k = ~0ull;
    if (kind == RFA_DISTANCE && is_int_kind(metric, kind) && metric != RFA_LCSSEQ) {
        if (metric == RFA_LEVENSHTEIN)
            s.band_k = s.ins ? host_ceil_div(cu, s.ins) : ~0ull;
        else
            s.band_k = cu;
    }
    return s;
}

ScoreSpec make_spec(const rfa_scorer* sc, int kind, int semantics, uint64_t cu, double cd, int algo)
{
    ScoreSpec s{};
    s.metric = sc->metric;
    s.kind = kind;
    s.semantics = semantics;
    s.algo = algo;
    s.cutoff_u = cu;
    s.cutoff_d = cd;
    s.empty_zero = sc->empty_zero ? 1 : 0;
    s.ins = sc->w.insert_cost;
    s.del = sc->w.delete_cost;
    s.rep = sc->w.replace_cost;
    s.prefix_weight = sc->prefix_weight;
    s.band_k = ~0ull;
    if (kind == RFA_DISTANCE && is_int_kind(sc->metric, kind) && sc->metric != RFA_LCSSEQ) {
        if (sc->metric == RFA_LEVENSHTEIN)
            s.band_k = s.ins ? host_ceil_div(cu, s.ins) : ~0ull;
        else
            s.band_k = cu;
    }
    return s;
}

// Whether the cutoff can decide a result from the two lengths alone.  The
// reference's defaults (no integer cutoff, similarity 0, normalized 1.0 / 0.0)
// never do, so those calls skip the per-pair test.
bool cutoff_may_settle(const ScoreSpec& s, bool per_query)
{
    if (per_query) return true;
    const bool fp = s.metric == RFA_JARO || s.metric == RFA_JARO_WINKLER || s.metric == RFA_RATIO;
    switch (s.kind) {
    case RFA_DISTANCE: return fp ? s.cutoff_d < 1.0 : s.cutoff_u < (1ull << 62);
    case RFA_SIMILARITY: return fp ? s.cutoff_d > 0.0 : s.cutoff_u > 0;
    case RFA_NORMALIZED_DISTANCE: return s.cutoff_d < 1.0;
    default: return s.cutoff_d > 0.0;
    }
}

void validate_kind(int metric, int kind, int semantics)
{
    if (kind < RFA_DISTANCE || kind > RFA_NORMALIZED_SIMILARITY) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown kind");
    if (semantics != RFA_CACHED && semantics != RFA_MULTI) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown semantics");
    if ((metric == RFA_RATIO || metric == RFA_PARTIAL_RATIO) && kind != RFA_SIMILARITY)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "ratio only provides similarity");
    if (semantics == RFA_MULTI && metric == RFA_PARTIAL_RATIO)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "partial_ratio has no Multi* scorer");
    if (semantics == RFA_MULTI && (metric == RFA_HAMMING || metric == RFA_PREFIX || metric == RFA_POSTFIX))
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "Hamming, Prefix and Postfix have no Multi* scorer");
    if (semantics == RFA_MULTI && metric == RFA_DAMERAU_LEVENSHTEIN)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "Damerau-Levenshtein has no Multi* scorer");
}

// Block-kernel launch over a cross product (queries x choices) or explicit items.
void run_block(rfa_context* ctx, int algo, int width, BlockArgs a, uint64_t max_pattern, uint64_t max_text,
               bool reuse, uint32_t* queue = nullptr, uint64_t min_pattern = 0, uint32_t u8_rows = 0)
{
    if (a.nitems == 0) return;
    hipStream_t s = ctx->stream;
    const uint32_t W = (uint32_t)((max_pattern + 63) / 64);
    const uint32_t slots = std::max<uint32_t>(1, (W + 63) / 64);
    if (slots > 32) throw ApiError(RFA_ERR_UNSUPPORTED, "patterns longer than 131072 code units are not supported");
    // patterns whose map + PM rows fit in LDS_BUDGET run from LDS; larger ones from a
    // global slice of the workgroup (sized for the largest pattern).  When not even the
    // shortest pattern of the launch fits, no LDS is reserved, so the LDS does not cap
    // the number of waves per CU.
    constexpr uint64_t LDS_BUDGET = 40 * 1024;
    const uint64_t need_max = block_arena_bytes(width, (uint32_t)max_pattern, u8_rows);
    const uint64_t need_min = block_arena_bytes(width, (uint32_t)min_pattern, u8_rows ? 2 : 0);
    a.lds_bytes = need_min > LDS_BUDGET ? 0u : (uint32_t)std::min<uint64_t>(need_max, LDS_BUDGET);
    a.reuse_pattern = reuse ? 1u : 0u;
    // (banded launches also hold an 8 KiB q-gram set; their register footprint, not the LDS,
    // bounds them to 3 waves per SIMD, and surplus persistent workgroups only start late)
    const uint32_t waves_per_cu =
        (uint32_t)std::min<uint64_t>(16, std::max<uint64_t>(1, ctx->lds_per_cu / (a.lds_bytes + 512)));
    uint32_t nblocks = (uint32_t)std::min<uint64_t>(a.nitems, (uint64_t)ctx->num_cus * waves_per_cu);
    if (need_max > LDS_BUDGET) {
        // global arenas: at most 2 GiB in total, at least one workgroup per CU
        constexpr uint64_t ARENA_TOTAL = 2ull << 30;
        const uint64_t cap = std::max<uint64_t>((uint64_t)ctx->num_cus, ARENA_TOTAL / need_max);
        nblocks = (uint32_t)std::min<uint64_t>(nblocks, cap);
        a.pm_scratch_words = (need_max + 7) / 8;
        a.pm_scratch = (uint64_t*)ctx->pm_scratch.get(a.pm_scratch_words * 8 * nblocks);
    }
    if (algo == ALGO_JARO) {
        a.tflag_words = (max_text + 63) / 64 + 1;
        a.tflag_scratch = (uint64_t*)ctx->tflag.get(a.tflag_words * 8 * nblocks);
    }
    if (queue) {
        a.queue = queue; // reset by the planning kernels on this stream
    }
    else {
        a.queue = (uint32_t*)ctx->queue.get(64);
        HIP_CHECK(hipMemsetAsync(a.queue, 0, 4, s));
    }
    // device-planned items are few and sorted most expensive first: one per grab;
    // cross products (every query x every choice, most settled by the length filter)
    // are taken in batches of consecutive items, about 8 grabs per workgroup
    a.batch = a.nitems_dev ? 1u : (uint32_t)std::min<uint64_t>(1024, std::max<uint64_t>(1, a.nitems / (8ull * nblocks)));
    HIP_CHECK(launch_block(algo, width, a, nblocks, slots, s));
}

// token_set_ratio / token_ratio over the raw choice set: the deduplicated joins and token
// tables of both sides, then the set kernel; token_ratio first writes the token-sort
// ratio (its own scorer) and the set kernel combines.
void cdist_tset(rfa_scorer* sc, const rfa_strings* raw, int kind, int semantics, double cd, void* out,
                int out_on_device)
{
    rfa_context* ctx = sc->ctx;
    hipStream_t s = ctx->stream;
    if (kind != RFA_SIMILARITY) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "token ratios only provide similarity");
    if (semantics != RFA_CACHED)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "token_set_ratio and token_ratio have no Multi* scorer");
    if (raw->width != sc->q->width)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
    if (!sc->block_queries.empty() && sc->block_max_len > TSET_LONG_MAX)
        throw ApiError(RFA_ERR_UNSUPPORTED,
                       "the token-set ratios take queries whose deduplicated token join has at most 4096 code units");
    const uint64_t nq = sc->q->count, nc = raw->count;
    if (nq == 0 || nc == 0) return;
    if (nc > 0xffffffffull || nq > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 strings");
    const size_t bytes = (size_t)nq * nc * 8;
    void* dout = out_on_device ? out : ctx->out_tmp.get(bytes);
    const rfa_strings* ch = scorer_choices(sc, raw);
    const TokenTable& ct = *sc->c_tokens;
    int mode = 0;
    const double* aux = nullptr;
    switch (sc->api_metric) {
    case RFA_TOKEN_RATIO:
        cdist_impl(sc->sort_sub, scorer_choices(sc->sort_sub, raw), RFA_SIMILARITY, RFA_CACHED, 0, cd, dout, 1);
        mode = 1;
        break;
    case RFA_PARTIAL_TOKEN_SET_RATIO:
        cdist_impl(sc->pr_dedup, ch, RFA_SIMILARITY, RFA_CACHED, 0, cd, dout, 1);
        mode = 2;
        break;
    case RFA_PARTIAL_TOKEN_RATIO: {
        const rfa_strings* sorted_ch = scorer_choices(sc->sort_sub, raw);
        cdist_impl(sc->sort_sub, sorted_ch, RFA_SIMILARITY, RFA_CACHED, 0, cd, dout, 1);
        // The deduplicated join equals the sorted join unless the string repeats a token
        // (then it is shorter), so the partial_ratio of the deduplicated joins differs from
        // the sorted joins' only in the rows of queries and the columns of choices that
        // repeat one: start from the sorted joins' values and compute just those.
        if (!sc->dup_queries_ready) {
            const rfa_strings* qs = sc->sort_sub->q;
            for (uint64_t q = 0; q < nq; ++q)
                if (sc->q->len(q) < qs->len(q)) sc->dup_queries.push_back((uint32_t)q);
            if (!sc->dup_queries.empty()) {
                std::vector<uint8_t> chars;
                std::vector<uint64_t> off{0};
                const int w = sc->q->width;
                for (uint32_t q : sc->dup_queries) {
                    const uint8_t* p = sc->q->h_chars.data() + sc->q->h_off[q] * w;
                    chars.insert(chars.end(), p, p + sc->q->len(q) * w);
                    off.push_back(off.back() + sc->q->len(q));
                }
                rfa_strings* sub = make_strings(ctx, chars.data(), w, off.data(), sc->dup_queries.size());
                try {
                    sc->pr_dedup_dupq = make_scorer(ctx, RFA_PARTIAL_RATIO, sub, sc->w, sc->prefix_weight);
                }
                catch (...) {
                    free_strings(sub);
                    throw;
                }
                free_strings(sub);
            }
            sc->dup_queries_ready = true;
        }
        std::vector<uint32_t> dup_choices;
        for (uint64_t c = 0; c < nc; ++c)
            if (ch->len(c) < sorted_ch->len(c)) dup_choices.push_back((uint32_t)c);
        double* d2 = (double*)ctx->aux_tmp.get(bytes);
        if (sc->dup_queries.size() * 4 > nq || dup_choices.size() * 4 > nc || ctx->pcut) {
            // repeats are common (or the cutoffs are per pair): the whole matrix
            cdist_impl(sc->pr_dedup, ch, RFA_SIMILARITY, RFA_CACHED, 0, cd, d2, 1);
        }
        else {
            HIP_CHECK(hipMemcpyAsync(d2, dout, bytes, hipMemcpyDeviceToDevice, s));
            if (!dup_choices.empty()) {
                // all queries x the choices that repeat a token
                std::vector<uint8_t> chars;
                std::vector<uint64_t> off{0};
                const int w = ch->width;
                for (uint32_t c : dup_choices) {
                    const uint8_t* p = ch->h_chars.data() + ch->h_off[c] * w;
                    chars.insert(chars.end(), p, p + ch->len(c) * w);
                    off.push_back(off.back() + ch->len(c));
                }
                rfa_strings* sub = make_strings(ctx, chars.data(), w, off.data(), dup_choices.size());
                try {
                    const size_t sb = (size_t)nq * dup_choices.size() * 8;
                    double* d3 = (double*)ctx->tset_cols.get(sb + dup_choices.size() * 4);
                    uint32_t* cols = (uint32_t*)((uint8_t*)d3 + sb);
                    HIP_CHECK(hipMemcpyAsync(cols, dup_choices.data(), dup_choices.size() * 4, hipMemcpyHostToDevice, s));
                    cdist_impl(sc->pr_dedup, sub, RFA_SIMILARITY, RFA_CACHED, 0, cd, d3, 1);
                    HIP_CHECK(launch_scatter_cols(d2, nc, d3, cols, dup_choices.size(), nq, s));
                    HIP_CHECK(hipStreamSynchronize(s)); // the subset set is freed next
                }
                catch (...) {
                    free_strings(sub);
                    throw;
                }
                free_strings(sub);
            }
            if (!sc->dup_queries.empty()) {
                // the queries that repeat a token x all choices, row by row
                const uint64_t nd = sc->dup_queries.size();
                double* d4 = (double*)ctx->tset_rows.get((size_t)nd * nc * 8);
                cdist_impl(sc->pr_dedup_dupq, ch, RFA_SIMILARITY, RFA_CACHED, 0, cd, d4, 1);
                if (!sc->d_dup_queries) {
                    HIP_CHECK(hipMalloc(&sc->d_dup_queries, nd * sizeof(uint32_t)));
                    HIP_CHECK(hipMemcpy(sc->d_dup_queries, sc->dup_queries.data(), nd * sizeof(uint32_t),
                                        hipMemcpyHostToDevice));
                }
                HIP_CHECK(launch_scatter_rows(d2, nc, d4, sc->d_dup_queries, nd, s));
            }
        }
        aux = d2;
        mode = 3;
        break;
    }
    default: break;
    }
    if (!sc->classes.empty()) {
        TsetArgs a{};
        TileArgs& ta = a.t;
        ta.qchars = sc->q->d_chars;
        ta.qoff = sc->q->d_off;
        ta.qorder = sc->d_tile_order;
        ta.nq = (uint32_t)sc->tile_order.size();
        ta.cchars = ch->d_chars;
        ta.coff = ch->d_off;
        ta.ncsel = (uint32_t)nc;
        ta.out_stride = nc;
        ta.hash_bits = sc->hash_bits;
        ta.spec = make_spec(sc, RFA_SIMILARITY, RFA_CACHED, 0, cd, ALGO_TSET);
        ta.out = dout;
        ta.pcut = ctx->pcut;
        a.q_set_off = sc->q_tokens->d_off;
        a.q_set_start = sc->q_tokens->d_start;
        a.q_set_len = sc->q_tokens->d_len;
        a.q_set_hash = sc->q_tokens->d_hash;
        a.c_set_off = ct.d_off;
        a.c_set_start = ct.d_start;
        a.c_set_len = ct.d_len;
        a.c_set_hash = ct.d_hash;
        a.mode = mode;
        a.aux = aux;
        for (const TileClass& tc : sc->classes) {
            ta.tile_base = tc.first_tile;
            ta.rows_cap = tc.rows_cap;
            uint32_t target_blocks = (uint32_t)ctx->num_cus * 12;
            uint32_t nbx = std::max<uint32_t>(1, (target_blocks + tc.ntiles - 1) / tc.ntiles);
            uint32_t cpb = std::max<uint32_t>(8, (ta.ncsel + nbx - 1) / nbx);
            cpb = ((cpb + 7) / 8) * 8; // whole rounds of the 8 waves
            ta.choices_per_block = cpb;
            nbx = (ta.ncsel + cpb - 1) / cpb;
            HIP_CHECK(launch_tset(sc->q->width, tc.word_bits, a, tc.ntiles, nbx, s));
        }
    }
    if (!sc->block_queries.empty()) {
        // the queries the tile plan does not take: one wave per (query, choice)
        TsetLongArgs la{};
        PartialLongArgs& pa = la.p;
        pa.qchars = sc->q->d_chars;
        pa.qoff = sc->q->d_off;
        pa.qlist = sc->d_block_queries;
        pa.nql = (uint32_t)sc->block_queries.size();
        pa.cchars = ch->d_chars;
        pa.coff = ch->d_off;
        pa.nc = (uint32_t)nc;
        pa.out_stride = nc;
        pa.out = (double*)dout;
        pa.cutoff = cd;
        pa.pcut = ctx->pcut;
        la.q_set_off = sc->q_tokens->d_off;
        la.q_set_start = sc->q_tokens->d_start;
        la.q_set_len = sc->q_tokens->d_len;
        la.q_set_hash = sc->q_tokens->d_hash;
        la.c_set_off = ct.d_off;
        la.c_set_start = ct.d_start;
        la.c_set_len = ct.d_len;
        la.c_set_hash = ct.d_hash;
        la.mode = mode;
        la.aux = aux;
        const int width = sc->q->width;
        constexpr uint64_t LDS_BUDGET = 40 * 1024;
        const uint64_t need = block_arena_bytes(width, (uint32_t)sc->block_max_len, sc->block_u8_rows);
        const uint64_t items = (uint64_t)pa.nql * nc;
        uint32_t nblocks;
        if (need <= LDS_BUDGET) {
            pa.lds_bytes = (uint32_t)need;
            const uint32_t waves_per_cu =
                (uint32_t)std::min<uint64_t>(16, std::max<uint64_t>(1, ctx->lds_per_cu / (need + 1024)));
            nblocks = (uint32_t)std::min<uint64_t>(items, (uint64_t)ctx->num_cus * waves_per_cu);
        }
        else {
            constexpr uint64_t ARENA_TOTAL = 2ull << 30;
            const uint64_t cap = std::max<uint64_t>((uint64_t)ctx->num_cus, ARENA_TOTAL / need);
            nblocks = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(items, (uint64_t)ctx->num_cus * 8), cap);
            pa.pm_scratch_words = (need + 7) / 8;
            pa.pm_scratch = (uint64_t*)ctx->pm_scratch.get(pa.pm_scratch_words * 8 * std::max<uint32_t>(1, nblocks));
        }
        if (items) HIP_CHECK(launch_tset_long(width, la, std::max<uint32_t>(1, nblocks), s));
    }
    if (!out_on_device) {
        HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}

// fuzz::WRatio over the raw choice set (fuzz_impl.hpp:838-876): the ratio, then the token
// ratio, partial ratio and partial token ratio of every pair as whole cdist launches that
// read the cutoffs raised per pair (rfa_wratio.hip computes them and the final maximum).
void cdist_wratio(rfa_scorer* sc, const rfa_strings* raw, int kind, int semantics, double cd, void* out,
                  int out_on_device)
{
    rfa_context* ctx = sc->ctx;
    hipStream_t s = ctx->stream;
    if (kind != RFA_SIMILARITY) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "WRatio only provides similarity");
    if (semantics != RFA_CACHED) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "WRatio has no Multi* scorer");
    const uint64_t nq = sc->q->count, nc = raw->count;
    if (nq == 0 || nc == 0) return;
    if (nc > 0xffffffffull || nq > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 strings");
    const size_t bytes = (size_t)nq * nc * 8;
    double* dout = (double*)(out_on_device ? out : ctx->out_tmp.get(bytes));
    if (cd > 100) {
        HIP_CHECK(hipMemsetAsync(dout, 0, bytes, s)); // the reference returns 0 for every pair
    }
    else {
        WRatioArgs a{};
        a.qoff = sc->q->d_off;
        a.coff = raw->d_off;
        a.nq = (uint32_t)nq;
        a.nc = (uint32_t)nc;
        a.cutoff = cd;
        a.R = (double*)ctx->wr_buf[0].get(bytes);
        a.C = (double*)ctx->wr_buf[1].get(bytes);
        double* P = (double*)ctx->wr_buf[2].get(bytes);
        a.out = dout;
        cdist_impl(sc->wr_sub[0], scorer_choices(sc->wr_sub[0], raw), RFA_SIMILARITY, RFA_CACHED, 0, cd, a.R, 1);
        HIP_CHECK(launch_wratio(0, a, (uint32_t)ctx->num_cus, s));
        struct PcutScope {
            rfa_context* c;
            ~PcutScope() { c->pcut = nullptr; }
        } scope{ctx};
        ctx->pcut = a.C;
        cdist_tset(sc->wr_sub[1], raw, RFA_SIMILARITY, RFA_CACHED, cd, dout, 1);
        cdist_impl(sc->wr_sub[2], scorer_choices(sc->wr_sub[2], raw), RFA_SIMILARITY, RFA_CACHED, 0, cd, P, 1);
        a.X = P;
        HIP_CHECK(launch_wratio(1, a, (uint32_t)ctx->num_cus, s));
        cdist_tset(sc->wr_sub[3], raw, RFA_SIMILARITY, RFA_CACHED, cd, P, 1);
        HIP_CHECK(launch_wratio(2, a, (uint32_t)ctx->num_cus, s));
    }
    if (!out_on_device) {
        HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}

// partial_ratio: tile-planned queries (<= 64 units) through the tile kernel, the
// rest one wave per pair.  The long path holds the shorter string of each pair
// in registers: up to PARTIAL_MAX_WORDS * 64 units.
void cdist_partial(rfa_scorer* sc, const rfa_strings* ch, double cd, double* dout)
{
    rfa_context* ctx = sc->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t nc = ch->count;
    const int width = sc->q->width;
    ScoreSpec spec{};
    spec.metric = RFA_PARTIAL_RATIO;
    spec.kind = RFA_SIMILARITY;
    spec.semantics = RFA_CACHED;
    spec.algo = ALGO_PARTIAL;
    spec.cutoff_d = cd;
    spec.band_k = ~0ull;
    if (!sc->block_queries.empty()) {
        const uint64_t needle = std::min<uint64_t>(sc->block_max_len, ch->max_len);
        if (needle > PBIG_MAX)
            throw ApiError(RFA_ERR_UNSUPPORTED, "partial_ratio of two strings longer than 4096 code units");
    }
    if (!sc->classes.empty()) {
        TileArgs ta{};
        ta.qchars = sc->q->d_chars;
        ta.qoff = sc->q->d_off;
        ta.qorder = sc->d_tile_order;
        ta.nq = (uint32_t)sc->tile_order.size();
        ta.cchars = ch->d_chars;
        ta.coff = ch->d_off;
        ta.corder = nullptr;
        ta.ncsel = (uint32_t)nc;
        ta.out_stride = nc;
        ta.hash_bits = sc->hash_bits;
        ta.spec = spec;
        ta.out = dout;
        ta.pcut = ctx->pcut;
        for (const TileClass& tc : sc->classes) {
            ta.tile_base = tc.first_tile;
            ta.rows_cap = tc.rows_cap;
            uint32_t target_blocks = (uint32_t)ctx->num_cus * 12;
            uint32_t nbx = std::max<uint32_t>(1, (target_blocks + tc.ntiles - 1) / tc.ntiles);
            uint32_t cpb = std::max<uint32_t>(32, (ta.ncsel + nbx - 1) / nbx);
            cpb = ((cpb + 31) / 32) * 32; // whole 4-choice runs for each of the 8 waves
            ta.choices_per_block = cpb;
            nbx = (ta.ncsel + cpb - 1) / cpb;
            HIP_CHECK(launch_ptile(width, tc.word_bits, ta, tc.ntiles, nbx, s));
        }
    }
    if (!sc->block_queries.empty()) {
        PartialLongArgs pa{};
        pa.qchars = sc->q->d_chars;
        pa.qoff = sc->q->d_off;
        pa.qlist = sc->d_block_queries;
        pa.nql = (uint32_t)sc->block_queries.size();
        pa.cchars = ch->d_chars;
        pa.coff = ch->d_off;
        pa.nc = (uint32_t)nc;
        pa.out_stride = nc;
        pa.out = dout;
        pa.cutoff = cd;
        pa.pcut = ctx->pcut;
        const uint32_t needle_words = (uint32_t)std::min<uint64_t>(
            PARTIAL_MAX_WORDS, std::max<uint64_t>(1, (std::min<uint64_t>(sc->block_max_len, ch->max_len) + 63) / 64));
        constexpr uint64_t LDS_BUDGET = 40 * 1024;
        const uint64_t need = block_arena_bytes(width, (uint32_t)sc->block_max_len, sc->block_u8_rows);
        const uint64_t items = (uint64_t)pa.nql * nc;
        uint32_t nblocks;
        if (need <= LDS_BUDGET) {
            pa.lds_bytes = (uint32_t)need;
            const uint32_t waves_per_cu =
                (uint32_t)std::min<uint64_t>(16, std::max<uint64_t>(1, ctx->lds_per_cu / (need + 512)));
            nblocks = (uint32_t)std::min<uint64_t>(items, (uint64_t)ctx->num_cus * waves_per_cu);
        }
        else {
            constexpr uint64_t ARENA_TOTAL = 2ull << 30;
            const uint64_t cap = std::max<uint64_t>((uint64_t)ctx->num_cus, ARENA_TOTAL / need);
            nblocks = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(items, (uint64_t)ctx->num_cus * 8), cap);
            pa.pm_scratch_words = (need + 7) / 8;
            pa.pm_scratch = (uint64_t*)ctx->pm_scratch.get(pa.pm_scratch_words * 8 * nblocks);
        }
        if (items) HIP_CHECK(launch_plong(width, needle_words, pa, std::max<uint32_t>(1, nblocks), s));
        if (items && std::min<uint64_t>(sc->block_max_len, ch->max_len) > PRALIGN_MAX) {
            // pairs whose shorter string has more than 1024 units: one wave per pair
            PbigArgs ba{};
            ba.achars = sc->q->d_chars;
            ba.aoff = sc->q->d_off;
            ba.bchars = ch->d_chars;
            ba.boff = ch->d_off;
            ba.items = sc->d_block_queries;
            ba.nb = (uint32_t)nc;
            ba.nitems = items;
            ba.out_stride = nc;
            ba.cdist = 1;
            ba.cutoff = cd;
            ba.pcut = ctx->pcut;
            ba.score_out = (double*)dout;
            const uint64_t l1max = std::min<uint64_t>(sc->block_max_len, ch->max_len);
            const uint64_t l2max = std::max<uint64_t>(sc->block_max_len, ch->max_len);
            ba.ws_stride = pbig_ws_bytes(width, l1max, l1max + l2max);
            const uint32_t nb = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(std::min<uint64_t>(items, (uint64_t)ctx->num_cus * 4), (1ull << 30) / ba.ws_stride));
            ba.ws = (uint8_t*)ctx->pb_ws.get(nb * ba.ws_stride);
            HIP_CHECK(launch_pbig(width, ba, nb, s));
        }
    }
}

// Levenshtein with general weights: tile-planned queries (<= 64 units) through the
// tile kernel, the rest one wave per pair.  The long kernel holds up to
// WLEV_LONG_MAX pattern rows across a wave: the query, or the choice when the
// query is longer; when both are longer, the query in strips of WLEV_LONG_MAX rows.
void cdist_wlev(rfa_scorer* sc, const rfa_strings* ch, const ScoreSpec& spec, void* dout, const uint64_t* d_qcut)
{
    rfa_context* ctx = sc->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t nc = ch->count;
    const int width = sc->q->width;
    // every value of the recurrence stays below (|s1| + |s2| + 1) * max weight, with the
    // replace weight clamped to insert + delete (a larger one is never used)
    constexpr uint64_t W24 = 1ull << 24;
    if (spec.ins >= W24 || spec.del >= W24 || std::min<uint64_t>(spec.rep, spec.ins + spec.del) >= W24)
        throw ApiError(RFA_ERR_UNSUPPORTED, "Levenshtein weights of 2^24 or more are not accelerated");
    const uint64_t rep = std::min<uint64_t>(spec.rep, spec.ins + spec.del);
    const uint64_t wmax = std::max<uint64_t>(std::max<uint64_t>(spec.ins, spec.del), rep);
    const uint64_t lens = sc->q->max_len + ch->max_len + 1;
    if (wmax && (lens >= (1ull << 30) || wmax * lens >= (1ull << 30)))
        throw ApiError(RFA_ERR_UNSUPPORTED,
                       "Levenshtein with general weights: weights times string lengths must stay below 2^30");
    if (!sc->classes.empty()) {
        TileArgs ta{};
        ta.qchars = sc->q->d_chars;
        ta.qoff = sc->q->d_off;
        ta.qorder = sc->d_tile_order;
        ta.nq = (uint32_t)sc->tile_order.size();
        ta.cchars = ch->d_chars;
        ta.coff = ch->d_off;
        ta.corder = nullptr;
        ta.ncsel = (uint32_t)nc;
        ta.out_stride = nc;
        ta.hash_bits = sc->hash_bits;
        ta.spec = spec;
        ta.out = dout;
        ta.qcut = d_qcut;
        for (const TileClass& tc : sc->classes) {
            ta.tile_base = tc.first_tile;
            ta.rows_cap = tc.rows_cap;
            uint32_t target_blocks = (uint32_t)ctx->num_cus * 12;
            uint32_t nbx = std::max<uint32_t>(1, (target_blocks + tc.ntiles - 1) / tc.ntiles);
            uint32_t cpb = std::max<uint32_t>(32, (ta.ncsel + nbx - 1) / nbx);
            cpb = ((cpb + 31) / 32) * 32; // whole 4-choice runs for each of the 8 waves
            ta.choices_per_block = cpb;
            nbx = (ta.ncsel + cpb - 1) / cpb;
            HIP_CHECK(launch_wtile(width, tc.word_bits, ta, tc.ntiles, nbx, s));
        }
    }
    if (!sc->block_queries.empty()) {
        // a query over WLEV_LONG_MAX units holds the choice across the lanes, or, when the
        // choice is that long too, runs in strips of WLEV_LONG_MAX rows
        const bool strips = sc->block_max_len > WLEV_LONG_MAX && ch->max_len > WLEV_LONG_MAX;
        const uint64_t pattern = std::min<uint64_t>(sc->block_max_len, WLEV_LONG_MAX);
        WlevLongArgs wa{};
        wa.qchars = sc->q->d_chars;
        wa.qoff = sc->q->d_off;
        wa.qlist = sc->d_block_queries;
        wa.nql = (uint32_t)sc->block_queries.size();
        wa.cchars = ch->d_chars;
        wa.coff = ch->d_off;
        wa.nc = (uint32_t)nc;
        wa.out_stride = nc;
        wa.spec = spec;
        wa.out = dout;
        wa.qcut = d_qcut;
        const uint32_t cells = (uint32_t)std::max<uint64_t>(1, (pattern + 63) / 64);
        const uint64_t items = (uint64_t)wa.nql * nc;
        uint32_t nblocks = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>((items + 3) / 4, (uint64_t)ctx->num_cus * 8));
        if (strips && items) {
            // one boundary row (the longest choice + 1 values) per wave, at most 1 GiB in all
            wa.bnd_stride = (ch->max_len + 1 + 63) / 64 * 64;
            const uint64_t per_block = 4 * wa.bnd_stride * sizeof(int32_t);
            nblocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nblocks, (1ull << 30) / per_block));
            wa.bnd = (int32_t*)ctx->wl_bnd.get(nblocks * per_block);
        }
        if (items) HIP_CHECK(launch_wlong(width, cells, wa, nblocks, s));
    }
}

// Damerau-Levenshtein: the tile kernel over the scorer's query tiles; queries the
// tile plan does not take (longer than 64 units, or tiles of too many distinct
// symbols) one lane per pair with the arrays in scratch.
void cdist_dl(rfa_scorer* sc, const rfa_strings* ch, const ScoreSpec& spec, void* dout, const uint64_t* d_qcut)
{
    rfa_context* ctx = sc->ctx;
    if (ch->max_len >= (1ull << 29) || sc->q->max_len >= (1ull << 29))
        throw ApiError(RFA_ERR_UNSUPPORTED, "Damerau-Levenshtein strings of 2^29 or more code units");
    if (!sc->block_queries.empty() && ch->count) {
        DlLongArgs la{};
        la.qchars = sc->q->d_chars;
        la.qoff = sc->q->d_off;
        la.qlist = sc->d_block_queries;
        la.nql = (uint32_t)sc->block_queries.size();
        la.max_rows = (uint32_t)sc->block_max_len;
        la.cchars = ch->d_chars;
        la.coff = ch->d_off;
        la.nc = ch->count;
        la.out_stride = ch->count;
        la.spec = spec;
        la.out = dout;
        la.qcut = d_qcut;
        const uint64_t per_block = 3ull * (la.max_rows + 1) * 64 * 4;
        // one-wave blocks, up to 8 per SIMD: the row loop waits on scratch, so it needs the waves
        constexpr uint64_t SCRATCH_TOTAL = 4ull << 30;
        const uint64_t items = (uint64_t)la.nql * ((ch->count + 63) / 64);
        const uint64_t nb = std::max<uint64_t>(
            1, std::min<uint64_t>(std::min<uint64_t>(items, (uint64_t)ctx->num_cus * 32), SCRATCH_TOTAL / per_block));
        la.scratch = (uint32_t*)ctx->pm_scratch.get(nb * per_block);
        HIP_CHECK(launch_dllong(sc->q->width, la, (uint32_t)nb, ctx->stream));
    }
    const uint64_t nc = ch->count;
    TileArgs ta{};
    ta.qchars = sc->q->d_chars;
    ta.qoff = sc->q->d_off;
    ta.qorder = sc->d_tile_order;
    ta.nq = (uint32_t)sc->tile_order.size();
    ta.cchars = ch->d_chars;
    ta.coff = ch->d_off;
    ta.corder = nullptr;
    ta.ncsel = (uint32_t)nc;
    ta.out_stride = nc;
    ta.hash_bits = sc->hash_bits;
    ta.spec = spec;
    ta.out = dout;
    ta.qcut = d_qcut;
    for (const TileClass& tc : sc->classes) {
        ta.tile_base = tc.first_tile;
        ta.rows_cap = tc.rows_cap;
        uint32_t target_blocks = (uint32_t)ctx->num_cus * 12;
        uint32_t nbx = std::max<uint32_t>(1, (target_blocks + tc.ntiles - 1) / tc.ntiles);
        uint32_t cpb = std::max<uint32_t>(32, (ta.ncsel + nbx - 1) / nbx);
        cpb = ((cpb + 31) / 32) * 32; // whole 4-choice runs for each of the 8 waves
        ta.choices_per_block = cpb;
        nbx = (ta.ncsel + cpb - 1) / cpb;
        HIP_CHECK(launch_dltile(sc->q->width, tc.word_bits, ta, tc.ntiles, nbx, ctx->stream));
    }
}

// Window-size buckets of the band kernel's launch settings: band_words(k) in {1}, {2, 3},
// {4 .. 7}, {8 .. 15}, {16 .. 31}, {32 .. 64}.
uint32_t band_bucket(uint32_t k)
{
    const uint32_t B = band_words(k);
    uint32_t b = 0;
    while (b < 5 && (2u << b) <= B) ++b;
    return b;
}

// Words per lane without a measurement: the choice that keeps more of a wave's word slots
// (64 lanes x words per lane) on window words a pair needs, one word per lane on a tie.
uint32_t band_rule_wpl(uint32_t k)
{
    const uint32_t B = band_words(k);
    if (B < 2) return 1;
    // window words per lane-step: one per lane (64 lanes) or two (128 slots)
    const uint32_t used1 = band_pairs_per_wave(k, 1) * B, used2 = band_pairs_per_wave(k, 2) * B;
    return used2 > 2 * used1 ? 2u : 1u;
}

// Long 8-bit queries under a narrow Levenshtein distance cutoff run the diagonal band
// kernel: Cached distance with uniform weights, band_words(k) <= 64 and the query's
// padded PM rows within BAND_LDS_MAX.  The other block queries are returned in `rest`.
// `only`: forced settings for the autotuner's runs -- the band queries of one bucket with
// words per lane `wpl` (bucket < 6), or all of them with the chosen words per lane (bucket 6);
// `threads` the workgroup size.
struct BandForce {
    uint32_t bucket, wpl, threads;
};

uint32_t run_band_queries(rfa_scorer* sc, const rfa_strings* ch, const ScoreSpec& spec, int algo, int kind,
                          int semantics, const uint64_t* qcut_host, const uint64_t* d_qcut, void* dout,
                          std::vector<uint32_t>& rest, const BandForce* only = nullptr)
{
    constexpr uint64_t BAND_LDS_MAX = 64 * 1024;
    rfa_context* ctx = sc->ctx;
    rest.clear();
    const bool kernel_ok = kind == RFA_DISTANCE && semantics == RFA_CACHED && algo == ALGO_LEV && sc->q->width == 1 &&
                           spec.ins > 0 && sc->q_rows.size() == sc->q->count && !ctx->no_band;
    // one launch; each query runs with its own words per lane
    std::vector<uint32_t> lq;
    std::vector<uint8_t> lwpl;
    uint64_t lds = 0;
    bool multi = false, wpl2 = false;
    const uint32_t threads = only ? only->threads : (ctx->band_threads ? ctx->band_threads : BAND_THREADS);
    for (uint32_t qi : sc->block_queries) {
        const uint64_t m = sc->q->len(qi);
        const uint64_t cut = qcut_host ? qcut_host[qi] : spec.cutoff_u;
        const uint64_t k = spec.ins ? host_ceil_div(cut, spec.ins) : ~0ull; // insert weight 0: no band
        bool ok = kernel_ok && k < m && band_words((uint32_t)std::min<uint64_t>(k, 1u << 20)) <= BAND_MAX_WORDS;
        uint32_t wpl = 1;
        if (ok) {
            const uint32_t b = band_bucket((uint32_t)k);
            if (only && only->bucket < 6) {
                if (b != only->bucket) continue; // another bucket: not part of this run
                wpl = only->wpl;
            }
            else {
                wpl = ctx->band_wpl[b] ? ctx->band_wpl[b] : band_rule_wpl((uint32_t)k);
            }
            ok = band_lds_bytes((uint32_t)m, (uint32_t)k, sc->q_rows[qi], wpl) <= BAND_LDS_MAX;
        }
        if (!ok) {
            if (!only) rest.push_back(qi);
            continue;
        }
        lq.push_back(qi);
        lwpl.push_back((uint8_t)wpl);
        lds = std::max(lds, band_lds_bytes((uint32_t)m, (uint32_t)k, sc->q_rows[qi], wpl));
        multi = multi || band_words((uint32_t)k) > 1;
        wpl2 = wpl2 || wpl == 2;
    }
    const uint32_t launched = (uint32_t)lq.size();
    if (ch->count == 0 || lq.empty()) return launched;
    {
        BandArgs a{};
        a.qchars = (const uint8_t*)sc->q->d_chars;
        a.qoff = sc->q->d_off;
        a.queries = upload(ctx->keep, ctx->band_q, lq, ctx->stream);
        a.wpl = upload(ctx->keep, ctx->band_w, lwpl, ctx->stream);
        a.cchars = (const uint8_t*)ch->d_chars;
        a.coff = ch->d_off;
        a.nc = (uint32_t)ch->count;
        // about 8 workgroups of 4 waves per CU over the launch, at least 256 choices each
        const uint64_t target = (uint64_t)ctx->num_cus * 8 * 256 / threads;
        const uint64_t per_q = std::max<uint64_t>(1, (target + lq.size() - 1) / lq.size());
        uint64_t chunk = std::max<uint64_t>(256, (ch->count + per_q - 1) / per_q);
        chunk = (chunk + 63) / 64 * 64;
        a.chunk = (uint32_t)chunk;
        a.chunks_per_query = (uint32_t)((ch->count + chunk - 1) / chunk);
        a.out_stride = ch->count;
        a.spec = spec;
        a.out = dout;
        a.qcut = d_qcut;
        a.counters = (unsigned long long*)ctx->d_counters;
        const uint64_t nblocks = (uint64_t)lq.size() * a.chunks_per_query;
        if (nblocks > 0x7fffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "too many banded query chunks in one call");
        HIP_CHECK(launch_band(a, (uint32_t)nblocks, (uint32_t)lds, multi, wpl2, threads, ctx->stream));
    }
    return launched;
}

void cdist_impl(rfa_scorer* sc, const rfa_strings* ch, int kind, int semantics, uint64_t cu, double cd, void* out,
                int out_on_device, const uint64_t* qcut_host)
{
    rfa_context* ctx = sc->ctx;
    validate_kind(sc->metric, kind, semantics);
    if (semantics == RFA_MULTI && sc->api_metric != sc->metric)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "token_sort_ratio, partial_token_sort_ratio and QRatio have no Multi* scorer");
    const uint64_t* d_qcut = nullptr;
    if (qcut_host) {
        if (!is_int_kind(sc->metric, kind))
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "per-query cutoffs are for the integer kinds");
        std::vector<uint64_t> v(qcut_host, qcut_host + sc->q->count);
        cu = v.empty() ? 0 : *std::max_element(v.begin(), v.end()); // launch-level bound; items use their own
        d_qcut = upload(ctx->keep, ctx->qcut, v, ctx->stream);
    }
    if (ch->width != sc->q->width)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
    int algo = algo_for(sc->metric, sc->w);
    if (semantics == RFA_MULTI && sc->metric == RFA_LEVENSHTEIN) {
        // MultiLevenshtein takes weights {1, 1, <= 2} (Levenshtein.hpp:334-340) and always runs
        // the uniform recurrence (Levenshtein.hpp:383-399); maximum() uses the weights.
        if (!multi_levenshtein_weights(sc->w)) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unsupported weights");
        algo = ALGO_LEV;
    }
    if (algo < 0) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown metric");
    hipStream_t s = ctx->stream;
    const uint64_t nq = sc->q->count, nc = ch->count;
    const size_t bytes = (size_t)nq * nc * 8;
    if (nq == 0 || nc == 0) return;
    if (nc > 0xffffffffull || nq > 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 strings");
    void* dout = out_on_device ? out : ctx->out_tmp.get(bytes);
    const ScoreSpec spec = make_spec(sc, kind, semantics, cu, cd, algo);
    const int width = sc->q->width;
    if (algo == ALGO_PARTIAL) {
        cdist_partial(sc, ch, cd, (double*)dout);
        if (!out_on_device) {
            HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return;
    }
    if (algo == ALGO_WLEV || algo == ALGO_DL) {
        if (algo == ALGO_WLEV)
            cdist_wlev(sc, ch, spec, dout, d_qcut);
        else
            cdist_dl(sc, ch, spec, dout, d_qcut);
        if (!out_on_device) {
            HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return;
    }
    if (algo == ALGO_POSIT) {
        // positional metrics: one kernel for every query length
        PositArgs pa{};
        pa.qchars = sc->q->d_chars;
        pa.qoff = sc->q->d_off;
        pa.nq = (uint32_t)nq;
        pa.cchars = ch->d_chars;
        pa.coff = ch->d_off;
        pa.nc = (uint32_t)nc;
        pa.spec = spec;
        pa.out = dout;
        pa.out_stride = nc;
        pa.qcut = d_qcut;
        HIP_CHECK(launch_posit(width, pa, (uint32_t)ctx->num_cus, s));
        if (!out_on_device) {
            HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return;
    }
    // per-pair cutoffs (WRatio's sub-scores): the LCS kinds, which apply them after an exact LCS
    if (ctx->pcut && (algo != ALGO_LCS || semantics != RFA_CACHED))
        throw ApiError(RFA_ERR_UNSUPPORTED, "per-pair cutoffs are for the ratio scorers");
    // a few queries against a set: one pair per lane (choices across lanes)
    if (!ctx->pcut && cdist_few_queries(sc, ch, spec, algo, dout, d_qcut)) {
        if (!out_on_device) {
            HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return;
    }

    // choices routed to the block kernel for Jaro (texts longer than one word)
    std::vector<uint32_t> jaro_long;
    std::vector<uint32_t> tile_choices;
    bool all_choices = true;
    if (algo == ALGO_JARO) {
        for (uint64_t c = 0; c < nc; ++c) {
            if (ch->len(c) > 64)
                jaro_long.push_back((uint32_t)c);
            else
                tile_choices.push_back((uint32_t)c);
        }
        all_choices = jaro_long.empty();
    }
    // ---- tile kernel ----
    if (!sc->classes.empty()) {
        TileArgs ta{};
        ta.qchars = sc->q->d_chars;
        ta.qoff = sc->q->d_off;
        ta.qorder = sc->d_tile_order;
        ta.nq = (uint32_t)sc->tile_order.size();
        ta.cchars = ch->d_chars;
        ta.coff = ch->d_off;
        ta.corder = all_choices ? nullptr : upload(ctx->keep, ctx->misc, tile_choices, s);
        ta.ncsel = all_choices ? (uint32_t)nc : (uint32_t)tile_choices.size();
        ta.out_stride = nc;
        ta.hash_bits = sc->hash_bits;
        ta.spec = spec;
        ta.out = dout;
        ta.qcut = d_qcut;
        ta.settle = cutoff_may_settle(spec, d_qcut != nullptr) && !ctx->pcut ? 1u : 0u;
        {
            // the fast normalized epilogue computes maxima in 32 bits
            NormEpi ne;
            ta.norm_epi = norm_epi_setup(spec, ne) && (uint64_t)64 + ch->max_len <= 0xffffffffull && !ctx->pcut ? 1u : 0u;
        }
        ta.pcut = ctx->pcut;
        {
            // a Cached integer result is a distance, similarity or cutoff + 1 no larger than the
            // maximum, (query + choice length) x the largest weight (Multi* similarities can
            // wrap around below zero, as the reference's do, so they keep 64 bits)
            const uint64_t wmax = std::max({spec.ins, spec.del, spec.rep, (uint64_t)1});
            const uint64_t len = 64 + ch->max_len;
            ta.narrow_int = semantics == RFA_CACHED && len < (0xffffffffull / wmax) ? 1u : 0u;
        }
        if (ta.ncsel) {
            for (const TileClass& tc : sc->classes) {
                ta.tile_base = tc.first_tile;
                ta.rows_cap = tc.rows_cap;
                // enough workgroups to fill the chip several times over
                uint32_t target_blocks = (uint32_t)ctx->num_cus * 12;
                uint32_t nbx = std::max<uint32_t>(1, (target_blocks + tc.ntiles - 1) / tc.ntiles);
                uint32_t cpb = std::max<uint32_t>(64, (ta.ncsel + nbx - 1) / nbx);
                cpb = ((cpb + 63) / 64) * 64; // whole 8-choice runs for each of the 8 waves
                ta.choices_per_block = cpb;
                nbx = (ta.ncsel + cpb - 1) / cpb;
                HIP_CHECK(launch_tile(algo, width, tc.word_bits, ta, tc.ntiles, nbx, s));
            }
        }
    }
    // ---- long (or large-alphabet) queries x all choices: the diagonal band kernel for
    // narrow distance cutoffs, the block kernel for the rest ----
    std::vector<uint32_t> rest_queries;
    if (!sc->block_queries.empty()) run_band_queries(sc, ch, spec, algo, kind, semantics, qcut_host, d_qcut, dout, rest_queries);
    if (!rest_queries.empty()) {
        BlockArgs ba{};
        ba.achars = sc->q->d_chars;
        ba.aoff = sc->q->d_off;
        ba.bchars = ch->d_chars;
        ba.boff = ch->d_off;
        ba.item_a = rest_queries.size() == sc->block_queries.size() ? sc->d_block_queries
                                                                    : upload(ctx->keep, ctx->items_a, rest_queries, s);
        std::vector<uint32_t> all(nc);
        std::iota(all.begin(), all.end(), 0u);
        ba.item_b = upload(ctx->keep, ctx->items_b, all, s);
        ba.cross_nb = (uint32_t)nc;
        ba.out_stride = nc;
        uint64_t ni = (uint64_t)rest_queries.size() * nc;
        if (ni > BLOCK_MAX_ITEMS) throw ApiError(RFA_ERR_UNSUPPORTED, "too many long-query pairs in one call");
        ba.nitems = (uint32_t)ni;
        ba.spec = spec;
        ba.out = dout;
        ba.qcut = d_qcut;
        ba.pcut = ctx->pcut; // indexed like out: query * nc + choice
        // the q-gram filter can reject pairs only where 7k + 7 < m
        if (kind == RFA_DISTANCE)
            for (uint32_t qi : rest_queries) {
                const uint64_t kq = qcut_host ? qcut_host[qi] : cu;
                if (kq < (1ull << 40) && 7 * kq + 7 < sc->q->len(qi)) {
                    ba.qgram = 1;
                    break;
                }
            }
        run_block(ctx, algo, width, ba, sc->block_max_len, ch->max_len, true, nullptr, sc->block_min_len,
                  sc->block_u8_rows);
    }
    // ---- Jaro: short queries x long choices ----
    if (!jaro_long.empty() && !sc->tile_order.empty()) {
        BlockArgs ba{};
        ba.achars = sc->q->d_chars;
        ba.aoff = sc->q->d_off;
        ba.bchars = ch->d_chars;
        ba.boff = ch->d_off;
        ba.item_a = sc->d_tile_order;
        ba.item_b = upload(ctx->keep, ctx->items_a, jaro_long, s);
        ba.cross_nb = (uint32_t)jaro_long.size();
        ba.out_stride = nc;
        uint64_t ni = (uint64_t)sc->tile_order.size() * jaro_long.size();
        if (ni > BLOCK_MAX_ITEMS) throw ApiError(RFA_ERR_UNSUPPORTED, "too many long-choice pairs in one call");
        ba.nitems = (uint32_t)ni;
        ba.spec = spec;
        ba.out = dout;
        run_block(ctx, algo, width, ba, 64, ch->max_len, true);
    }
    if (!out_on_device) {
        HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}


// ---- independent pairs (the uneven batch) ---------------------------------

uint64_t* lane_arena(rfa_context* ctx, uint32_t maxw, uint32_t nlanes, uint64_t nitems, uint32_t* gen_base)
{
    const size_t words = lane_arena_words(maxw);
    const size_t bytes = words * 8 * (size_t)nlanes;
    uint64_t* arena = (uint64_t*)ctx->lane_arena.get(bytes);
    // generation tags must never repeat against stale slots: zero fresh memory,
    // and restart the counter (with a fresh zeroing) before it would wrap.
    const uint64_t per_lane = (nitems + nlanes - 1) / nlanes + 1;
    if (ctx->lane_arena_bytes_zeroed < bytes || (uint64_t)ctx->gen_next + per_lane >= 0xffffff00ull) {
        HIP_CHECK(hipMemsetAsync(arena, 0, ctx->lane_arena.cap, ctx->stream));
        ctx->lane_arena_bytes_zeroed = ctx->lane_arena.cap;
        ctx->gen_next = 1;
    }
    *gen_base = ctx->gen_next;
    ctx->gen_next += (uint32_t)per_lane;
    return arena;
}

// The bucketed pairs schedule after the metric upload: the device plan, then the
// segmented kernel (and the short-OSA lane kernels, and the block kernel with its
// queue on the first stream when some pattern needs it).  Every buffer is the
// caller's; the function allocates nothing unless need_block (the block kernel's
// pattern arenas).
static void pairs_bucketed_launch(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const uint8_t* dmet,
                                  uint64_t n, const ScoreSpec& base, bool has_osa, bool need_block, uint64_t maxm,
                                  uint64_t maxn, uint32_t nb, uint32_t* bins, uint32_t* vals, void* temp,
                                  size_t temp_bytes, uint32_t* counters, void* dout, int kind, uint64_t cu)
{
    hipStream_t s = ctx->stream, s2 = ctx->stream2;
    const int width = a->width;
    const uint32_t* packed = vals + n;
    LaneArgs la{};
    la.achars = a->d_chars;
    la.aoff = a->d_off;
    la.bchars = b->d_chars;
    la.boff = b->d_off;
    la.spec = base;
    la.out = dout;
    HIP_CHECK(launch_plan(a->d_off, b->d_off, dmet, n, ctx->seg_chain_steps, bins, vals, vals + n, temp,
                          temp_bytes, counters, s));
    // with block work, the segmented and lane kernels (second stream) overlap the block
    // queue (first stream); without, everything stays on the first stream
    hipStream_t sk = s;
    if (need_block) {
        HIP_CHECK(hipEventRecord(ctx->fork, s));
        HIP_CHECK(hipStreamWaitEvent(s2, ctx->fork, 0));
        sk = s2;
        BlockArgs ba{};
        ba.achars = a->d_chars;
        ba.aoff = a->d_off;
        ba.bchars = b->d_chars;
        ba.boff = b->d_off;
        ba.item_packed = packed;
        ba.nitems_dev = counters + 1;
        ba.pair_metric = dmet;
        ba.nitems = (uint32_t)n; // upper bound, for the grid size
        ba.spec = base;
        ba.out = dout;
        ba.qgram = kind == RFA_DISTANCE && cu < (1ull << 40) && 7 * cu + 7 < maxn ? 1u : 0u;
        run_block(ctx, ALGO_MIXED, width, ba, maxm, maxn, false, counters);
    }
    SegArgs sa{};
    sa.achars = a->d_chars;
    sa.aoff = a->d_off;
    sa.bchars = b->d_chars;
    sa.boff = b->d_off;
    sa.item_packed = packed;
    sa.pair_metric = dmet;
    sa.counters = counters;
    sa.spec = base;
    sa.out = dout;
    // no more waves than pairs (each wave holds up to 64 pairs, at least one)
    const uint32_t max_blocks = (uint32_t)std::min<uint64_t>((n + 3) / 4, 1u << 20);
    HIP_CHECK(launch_seg(width, nb, sa, (uint32_t)ctx->num_cus, max_blocks, sk));
    if (has_osa) {
        la.item_packed = packed;
        la.pair_metric = dmet;
        la.nitems = n;
        // OSA 33-64 unit range [counters[1], counters[2]), then 1-32 units [counters[2], counters[3])
        const uint32_t nblk = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * 6, (n + 63) / 64);
        la.begin_dev = counters + 1;
        la.end_dev = counters + 2;
        HIP_CHECK(launch_lane_short(width, 64, la, nblk, sk));
        la.begin_dev = counters + 2;
        la.end_dev = counters + 3;
        HIP_CHECK(launch_lane_short(width, 32, la, nblk, sk));
    }
    if (need_block) {
        HIP_CHECK(hipEventRecord(ctx->join, s2));
        HIP_CHECK(hipStreamWaitEvent(s, ctx->join, 0));
    }
}

void pairs_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const int32_t* metrics, int kind,
                uint64_t cu, double cd, int schedule, void* out, int out_on_device)
{
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    if (kind < RFA_DISTANCE || kind > RFA_NORMALIZED_SIMILARITY) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown kind");
    if (schedule != RFA_SCHEDULE_BUCKETED && schedule != RFA_SCHEDULE_PLAIN)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown schedule");
    const uint64_t n = a->count;
    if (n == 0) return;
    if (!metrics) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "metrics is NULL");
    if (n > BLOCK_MAX_ITEMS) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 - 2^26 pairs");
    // one pass: validate, note OSA, stage into pinned memory
    int slot_idx = 0;
    uint8_t* staged = (uint8_t*)ctx->staging.acquire(n, &slot_idx); // one byte per pair on the device
    bool has_osa = false;
    bool bad_metric = false;
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t m = metrics[i];
        bad_metric |= (m < RFA_LEVENSHTEIN) | (m > RFA_OSA);
        has_osa |= m == RFA_OSA;
        staged[i] = (uint8_t)m;
    }
    if (bad_metric) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "pair metrics must be Levenshtein, Indel, LCSseq or OSA");
    hipStream_t s = ctx->stream;
    const size_t bytes = (size_t)n * 8;
    void* dout = out_on_device ? out : ctx->out_tmp.get(bytes);
    const int width = a->width;
    const rfa_weights unit{1, 1, 1};
    ScoreSpec base = make_spec_m(RFA_LEVENSHTEIN, unit, 0.1, kind, RFA_CACHED, cu, cd, ALGO_LEV);

    LaneArgs la{};
    la.achars = a->d_chars;
    la.aoff = a->d_off;
    la.bchars = b->d_chars;
    la.boff = b->d_off;
    la.spec = base;
    la.out = dout;

    if (schedule == RFA_SCHEDULE_PLAIN) {
        // one pair per lane, input order, all words in the lane
        uint64_t maxm = 0;
        for (uint64_t i = 0; i < n; ++i) maxm = std::max(maxm, a->len(i));
        if (maxm > 4096) throw ApiError(RFA_ERR_UNSUPPORTED, "the plain schedule handles patterns up to 4096 units");
        uint8_t* dmet = (uint8_t*)ctx->lane_metric.get(n);
        ctx->staging.copy(slot_idx, dmet, n, s);
        la.item_metric = dmet;
        la.nitems = n;
        const uint32_t nlanes = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * 4 * 64,
                                                             ((n + 255) / 256) * 256);
        la.arena_words = lane_arena_words(64);
        la.arena = lane_arena(ctx, 64, nlanes, n, &la.gen_base);
        HIP_CHECK(launch_lane(width, 64, la, nlanes / 256, s));
    }
    else {
        // bucketed, planned on the device: pattern = the shorter string; Levenshtein / Indel /
        // LCSseq patterns of <= 4096 units -> segmented kernel, OSA patterns of <= 64 units ->
        // lane kernel, the rest -> block kernel (longest-processing-time-first queue)
        uint8_t* dmet = (uint8_t*)ctx->lane_metric.get(n);
        uint32_t* bins = (uint32_t*)ctx->plan_keys.get(n * 4);
        uint32_t* vals = (uint32_t*)ctx->plan_vals.get(n * 8);
        const size_t temp_bytes = plan_temp_bytes(n);
        void* temp = ctx->plan_temp.get(temp_bytes);
        uint32_t* counters = (uint32_t*)ctx->queue.get(128); // 18 range ends and the block queue, written by the plan
        const uint64_t maxm = std::min(a->max_len, b->max_len);
        const uint64_t maxn = std::max(a->max_len, b->max_len);
        const bool need_block = maxm > 4096 || (has_osa && maxm > 64);
        const uint32_t nb = seg_symbol_bits(width, std::max(a->max_unit(), b->max_unit()));
        // the launch sequence after the metric upload; everything it reads on the host is in
        // the graph key below
        auto enqueue = [&]() {
            pairs_bucketed_launch(ctx, a, b, dmet, n, base, has_osa, need_block, maxm, maxn, nb, bins, vals, temp,
                                  temp_bytes, counters, dout, kind, cu);
        };
        // every host value the launch sequence bakes in: same key, same launches
        auto key_now = [&]() {
            uint64_t cdb;
            std::memcpy(&cdb, &cd, 8);
            return std::vector<uint64_t>{(uint64_t)(uintptr_t)a->d_chars, (uint64_t)(uintptr_t)a->d_off,
                                         (uint64_t)(uintptr_t)b->d_chars, (uint64_t)(uintptr_t)b->d_off, n,
                                         (uint64_t)kind, cu, cdb, (uint64_t)(uintptr_t)dout, (uint64_t)width,
                                         (uint64_t)has_osa, maxm, maxn, nb, (uint64_t)ctx->num_cus,
                                         (uint64_t)ctx->seg_chain_steps,
                                         (uint64_t)(uintptr_t)ctx->lane_metric.p, ctx->lane_metric.cap,
                                         (uint64_t)(uintptr_t)ctx->plan_keys.p, ctx->plan_keys.cap,
                                         (uint64_t)(uintptr_t)ctx->plan_vals.p, ctx->plan_vals.cap,
                                         (uint64_t)(uintptr_t)ctx->plan_temp.p, ctx->plan_temp.cap,
                                         (uint64_t)(uintptr_t)ctx->queue.p, ctx->queue.cap,
                                         (uint64_t)(uintptr_t)ctx->pm_scratch.p, ctx->pm_scratch.cap};
        };
        // the metric upload stays outside the graph: its pinned source alternates per call
        ctx->staging.copy(slot_idx, dmet, n, s);
        PairsGraph& g = ctx->pairs_graph;
        const std::vector<uint64_t> key = ctx->pairs_graph_on ? key_now() : std::vector<uint64_t>{};
        if (!ctx->pairs_graph_on) {
            enqueue();
        }
        else if (g.exec && key == g.key) {
            HIP_CHECK(hipGraphLaunch(g.exec, s));
        }
        else if (key == g.last_eager) {
            // the same call as the last one, whose buffers are all sized: capture it once and
            // replay the graph from now on
            g.release();
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipGraph_t graph = nullptr;
            try {
                enqueue();
            }
            catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(s, &graph));
            const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) {
                g.exec = nullptr;
                throw HipError(ie, "hipGraphInstantiate");
            }
            g.key = key;
            HIP_CHECK(hipGraphLaunch(g.exec, s));
        }
        else {
            enqueue();
            g.last_eager = key_now();
        }
    }
    if (!out_on_device) {
        HIP_CHECK(hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}

// rfa_pairs_noalloc: the bucketed schedule of pairs whose shorter string has at
// most 4096 units (64 when OSA codes may occur), with the plan in this device's
// static workspace.  Allocates nothing, copies nothing, does not synchronize.
void pairs_noalloc_impl(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const uint8_t* d_metrics,
                        int has_osa, int kind, uint64_t cu, double cd, void* d_out)
{
    if (a->count != b->count) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must hold the same number of strings");
    if (a->width != b->width) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "a and b must have the same char width");
    if (kind < RFA_DISTANCE || kind > RFA_NORMALIZED_SIMILARITY) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown kind");
    const uint64_t n = a->count;
    if (n == 0) return;
    if (!d_metrics || !d_out) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "metrics or out is NULL");
    if (n > NOALLOC_MAX_PAIRS)
        throw ApiError(RFA_ERR_UNSUPPORTED, "rfa_pairs_noalloc takes at most 2^20 pairs per call");
    if (!ctx->noalloc_scratch) throw ApiError(RFA_ERR_NO_DEVICE, "the static workspace has no device address");
    // every pair must run in the segmented kernel (or, for OSA, the short lane kernel)
    const uint64_t cap = has_osa ? 64 : 4096;
    if (std::min(a->max_len, b->max_len) > cap)
        for (uint64_t i = 0; i < n; ++i)
            if (std::min(a->len(i), b->len(i)) > cap)
                throw ApiError(RFA_ERR_UNSUPPORTED, has_osa ? "rfa_pairs_noalloc with OSA takes pairs whose shorter "
                                                              "string has at most 64 units"
                                                            : "rfa_pairs_noalloc takes pairs whose shorter string has "
                                                              "at most 4096 units");
    // workspace: counters | bins (4 B per pair) | items (8 B per pair) | plan temporaries
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t temp_bytes = plan_temp_bytes(n);
    const size_t off_bins = 256, off_vals = off_bins + align(n * 4), off_temp = off_vals + align(n * 8);
    if (off_temp + temp_bytes > NOALLOC_SCRATCH_BYTES)
        throw ApiError(RFA_ERR_UNSUPPORTED, "the plan does not fit the static workspace");
    uint8_t* ws = (uint8_t*)ctx->noalloc_scratch;
    const rfa_weights unit{1, 1, 1};
    const ScoreSpec base = make_spec_m(RFA_LEVENSHTEIN, unit, 0.1, kind, RFA_CACHED, cu, cd, ALGO_LEV);
    const uint64_t maxm = std::min(a->max_len, b->max_len), maxn = std::max(a->max_len, b->max_len);
    const uint32_t nb = seg_symbol_bits(a->width, std::max(a->max_unit(), b->max_unit()));
    pairs_bucketed_launch(ctx, a, b, d_metrics, n, base, has_osa != 0, false, maxm, maxn, nb,
                          (uint32_t*)(ws + off_bins), (uint32_t*)(ws + off_vals), ws + off_temp, temp_bytes,
                          (uint32_t*)ws, d_out, kind, cu);
}

} // namespace rfa

using namespace rfa;

extern "C" {

int rfa_abi_version(void) { return RFA_ABI_VERSION; }

const char* rfa_last_error(void) { return g_err.c_str(); }

rfa_status rfa_context_create(int device, rfa_context** out)
{
    if (!out) return fail(RFA_ERR_INVALID_ARGUMENT, "out is NULL");
    return guarded([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            (void)hipGetLastError();
            throw ApiError(RFA_ERR_NO_DEVICE, "no HIP device available");
        }
        if (device < 0 || device >= n) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "device index out of range");
        HIP_CHECK(hipSetDevice(device));
        // the static workspace of rfa_pairs_noalloc: its address on this device (now current)
        void* ws = nullptr;
        HIP_CHECK(noalloc_scratch_address(device, &ws));
        auto* c = new rfa_context();
        c->noalloc_scratch = ws;
        c->device = device;
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        if (prop.maxSharedMemoryPerMultiProcessor > 0) c->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
        *out = c;
    });
}

rfa_status rfa_context_destroy(rfa_context* ctx)
{
    if (!ctx) return RFA_OK;
    return guarded([&] {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamSynchronize(ctx->stream2);
        for (DevBuf* b : {&ctx->queue, &ctx->pm_scratch, &ctx->tflag, &ctx->items_a, &ctx->items_b, &ctx->items_idx,
                          &ctx->items_swap, &ctx->out_tmp, &ctx->aux_tmp, &ctx->misc, &ctx->lane_arena, &ctx->lane_items,
                          &ctx->lane_swap, &ctx->lane_metric, &ctx->lane_out, &ctx->block_items[0],
                          &ctx->block_items[1], &ctx->block_items[2], &ctx->block_items[3], &ctx->plan_keys,
                          &ctx->plan_vals, &ctx->plan_temp, &ctx->topk_scores, &ctx->topk_run[0],
                          &ctx->topk_run[1], &ctx->topk_cand_s, &ctx->topk_cand_i, &ctx->topk_counts, &ctx->topk_q,
                          &ctx->topk_out, &ctx->qcut, &ctx->ed_items, &ctx->ed_rows, &ctx->ed_ops, &ctx->tset_cols,
                          &ctx->tset_rows, &ctx->band_q, &ctx->band_w, &ctx->wl_bnd, &ctx->ed_long, &ctx->ed_hnodes, &ctx->ed_htmp, &ctx->pb_ws, &ctx->wr_buf[0],
                          &ctx->wr_buf[1], &ctx->wr_buf[2]})
            b->release();
        ctx->staging.release();
        ctx->pairs_graph.release();
        if (ctx->d_counters) (void)hipFree(ctx->d_counters);
        (void)hipEventDestroy(ctx->fork);
        (void)hipEventDestroy(ctx->join);
        (void)hipStreamDestroy(ctx->stream2);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

rfa_status rfa_context_synchronize(rfa_context* ctx)
{
    if (!ctx) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    return guarded([&] { HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

void* rfa_context_stream(rfa_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

rfa_status rfa_context_counter(rfa_context* ctx, int counter, uint64_t* value)
{
    if (!ctx || !value) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx or value is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (counter != RFA_COUNTER_BAND_WORD_STEPS && counter != RFA_COUNTER_BAND_PAIRS)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown counter");
        *value = 0;
        HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->d_counters)
            HIP_CHECK(hipMemcpy(value, ctx->d_counters + (counter - 1), sizeof(uint64_t), hipMemcpyDeviceToHost));
    });
}

rfa_status rfa_context_set_option(rfa_context* ctx, int option, int64_t value)
{
    if (!ctx) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (option == RFA_OPTION_PAIRS_CHAIN) {
            if (value < 1 || value > (int64_t)1 << 31)
                throw ApiError(RFA_ERR_INVALID_ARGUMENT, "RFA_OPTION_PAIRS_CHAIN must be in [1, 2^31]");
            ctx->seg_chain_steps = (uint32_t)value;
            return;
        }
        if (option == RFA_OPTION_COUNT_WORD_STEPS) {
            if (value != 0 && value != 1) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "RFA_OPTION_COUNT_WORD_STEPS is 0 or 1");
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            if (value && !ctx->d_counters) HIP_CHECK(hipMalloc(&ctx->d_counters, 8 * sizeof(uint64_t)));
            if (ctx->d_counters) HIP_CHECK(hipMemset(ctx->d_counters, 0, 8 * sizeof(uint64_t)));
            if (!value && ctx->d_counters) {
                HIP_CHECK(hipFree(ctx->d_counters));
                ctx->d_counters = nullptr;
            }
            return;
        }
        if (option == RFA_OPTION_CDIST_BAND) {
            if (value != 0 && value != 1) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "RFA_OPTION_CDIST_BAND is 0 or 1");
            ctx->no_band = value == 0;
            return;
        }
        if (option != RFA_OPTION_PAIRS_GRAPH) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown option");
        if (value != 0 && value != 1) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "RFA_OPTION_PAIRS_GRAPH is 0 or 1");
        ctx->pairs_graph_on = value != 0;
        if (!ctx->pairs_graph_on) {
            HIP_CHECK(hipStreamSynchronize(ctx->stream));
            ctx->pairs_graph.release();
            ctx->pairs_graph.last_eager.clear();
        }
    });
}

rfa_status rfa_strings_create(rfa_context* ctx, const void* chars, int char_width, const uint64_t* offsets,
                              uint64_t count, rfa_strings** out)
{
    if (!ctx || !out) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx or out is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        *out = make_strings(ctx, chars, char_width, offsets, count);
    });
}

rfa_status rfa_strings_destroy(rfa_strings* set)
{
    if (!set) return RFA_OK;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(set->ctx->mu);
        (void)hipSetDevice(set->ctx->device);
        (void)hipStreamSynchronize(set->ctx->stream);
        free_strings(set);
    });
}

uint64_t rfa_strings_count(const rfa_strings* set) { return set ? set->count : 0; }

rfa_status rfa_scorer_create(rfa_context* ctx, int metric, const rfa_strings* queries, const rfa_weights* weights,
                             double prefix_weight, rfa_scorer** out)
{
    if (!ctx || !queries || !out) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, queries or out is NULL");
    return guarded([&] {
        if (metric < RFA_LEVENSHTEIN || metric > RFA_WRATIO)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown metric");
        rfa_weights w = weights ? *weights : rfa_weights{1, 1, 1};
        if (metric != RFA_LEVENSHTEIN) w = rfa_weights{1, 1, 1};
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        *out = make_scorer(ctx, metric, queries, w, prefix_weight);
    });
}

rfa_status rfa_scorer_set_option(rfa_scorer* sc, int option, int64_t value)
{
    if (!sc) return fail(RFA_ERR_INVALID_ARGUMENT, "scorer is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(sc->ctx->mu);
        HIP_CHECK(hipSetDevice(sc->ctx->device));
        if (option != RFA_SCORER_SIGNED_UNITS) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown scorer option");
        if (value != 0 && value != 1) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "RFA_SCORER_SIGNED_UNITS is 0 or 1");
        begin_call(sc->ctx);
        scorer_set_signed(sc, value == 1);
    });
}

rfa_status rfa_scorer_destroy(rfa_scorer* sc)
{
    if (!sc) return RFA_OK;
    return guarded([&] {
        std::lock_guard<std::mutex> lk(sc->ctx->mu);
        (void)hipSetDevice(sc->ctx->device);
        (void)hipStreamSynchronize(sc->ctx->stream);
        free_scorer(sc);
    });
}

rfa_status rfa_cdist(rfa_scorer* sc, const rfa_strings* choices, int kind, int semantics, uint64_t cutoff_u,
                     double cutoff_d, void* out, int out_on_device)
{
    if (!sc || !choices) return fail(RFA_ERR_INVALID_ARGUMENT, "scorer or choices is NULL");
    if (!out && sc->q->count && choices->count) return fail(RFA_ERR_INVALID_ARGUMENT, "out is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(sc->ctx->mu);
        HIP_CHECK(hipSetDevice(sc->ctx->device));
        begin_call(sc->ctx);
        if (choices->width != sc->q->width)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
        if (sc->wratio)
            cdist_wratio(sc, choices, kind, semantics, cutoff_d, out, out_on_device);
        else if (sc->token_set)
            cdist_tset(sc, choices, kind, semantics, cutoff_d, out, out_on_device);
        else
            cdist_impl(sc, scorer_choices(sc, choices), kind, semantics, cutoff_u, cutoff_d, out, out_on_device);
    });
}

rfa_status rfa_cdist_cutoffs(rfa_scorer* sc, const rfa_strings* choices, int kind, int semantics,
                             const uint64_t* cutoffs_u, void* out, int out_on_device)
{
    if (!sc || !choices) return fail(RFA_ERR_INVALID_ARGUMENT, "scorer or choices is NULL");
    if (!cutoffs_u && sc->q->count) return fail(RFA_ERR_INVALID_ARGUMENT, "cutoffs_u is NULL");
    if (!out && sc->q->count && choices->count) return fail(RFA_ERR_INVALID_ARGUMENT, "out is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(sc->ctx->mu);
        HIP_CHECK(hipSetDevice(sc->ctx->device));
        begin_call(sc->ctx);
        if (choices->width != sc->q->width)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
        cdist_impl(sc, scorer_choices(sc, choices), kind, semantics, 0, 0.0, out, out_on_device, cutoffs_u);
    });
}

rfa_status rfa_pairs(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const int32_t* metrics, int kind,
                     uint64_t cutoff_u, double cutoff_d, int schedule, void* out, int out_on_device)
{
    if (!ctx || !a || !b) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, a or b is NULL");
    if (!out && a->count) return fail(RFA_ERR_INVALID_ARGUMENT, "out is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        begin_call(ctx);
        pairs_impl(ctx, a, b, metrics, kind, cutoff_u, cutoff_d, schedule, out, out_on_device);
    });
}


rfa_status rfa_pairs_noalloc(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const uint8_t* metrics,
                             int has_osa, int kind, uint64_t cutoff_u, double cutoff_d, void* out)
{
    if (!ctx || !a || !b) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, a or b is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        pairs_noalloc_impl(ctx, a, b, metrics, has_osa, kind, cutoff_u, cutoff_d, out);
    });
}

rfa_status rfa_pairs_autotune(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const int32_t* metrics,
                              int kind, uint64_t cutoff_u, double cutoff_d, uint32_t* chosen)
{
    if (!ctx || !a || !b) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, a or b is NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        begin_call(ctx);
        // candidates for the length from which the longer string is the pattern and from
        // which pairs are scheduled first; 4097 keeps the shorter string everywhere
        static const uint32_t cand[] = {128, 256, 512, 1024, 2048, 4097};
        const uint32_t keep = ctx->seg_chain_steps;
        void* dout = ctx->out_tmp.get((size_t)a->count * 8 + 8);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        uint32_t best = keep;
        float best_ms = -1.0f;
        try {
            for (uint32_t c : cand) {
                ctx->seg_chain_steps = c;
                pairs_impl(ctx, a, b, metrics, kind, cutoff_u, cutoff_d, RFA_SCHEDULE_BUCKETED, dout, 1); // warm-up
                HIP_CHECK(hipEventRecord(e0, ctx->stream));
                for (int r = 0; r < 3; ++r)
                    pairs_impl(ctx, a, b, metrics, kind, cutoff_u, cutoff_d, RFA_SCHEDULE_BUCKETED, dout, 1);
                HIP_CHECK(hipEventRecord(e1, ctx->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (best_ms < 0.0f || ms < best_ms) {
                    best_ms = ms;
                    best = c;
                }
            }
        }
        catch (...) {
            ctx->seg_chain_steps = keep;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->seg_chain_steps = best;
        if (chosen) *chosen = best;
    });
}

rfa_status rfa_cdist_autotune(rfa_scorer* sc, const rfa_strings* choices, int kind, uint64_t cutoff_u,
                              const uint64_t* cutoffs_u, uint32_t* chosen)
{
    if (!sc || !choices) return fail(RFA_ERR_INVALID_ARGUMENT, "scorer or choices is NULL");
    return guarded([&] {
        rfa_context* ctx = sc->ctx;
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_CHECK(hipSetDevice(ctx->device));
        begin_call(ctx);
        if (chosen) std::fill(chosen, chosen + 12, 0u);
        validate_kind(sc->metric, kind, RFA_CACHED);
        if (choices->width != sc->q->width)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
        const int algo = algo_for(sc->metric, sc->w);
        const uint64_t nq = sc->q->count, nc = choices->count;
        if (nq == 0 || nc == 0 || sc->block_queries.empty()) return;
        const uint64_t* d_qcut = nullptr;
        uint64_t cu = cutoff_u;
        if (cutoffs_u) {
            std::vector<uint64_t> v(cutoffs_u, cutoffs_u + nq);
            cu = *std::max_element(v.begin(), v.end());
            d_qcut = upload(ctx->keep, ctx->qcut, v, ctx->stream);
        }
        const ScoreSpec spec = make_spec(sc, kind, RFA_CACHED, cu, 0.0, algo);
        void* dout = ctx->out_tmp.get((size_t)nq * nc * 8);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        std::vector<uint32_t> rest;
        try {
            // 1 warm-up and 2 timed runs of a forced setting; < 0 when no band query runs
            auto time_runs = [&](const BandForce& f) -> float {
                if (!run_band_queries(sc, choices, spec, algo, kind, RFA_CACHED, cutoffs_u, d_qcut, dout, rest, &f))
                    return -1.0f;
                HIP_CHECK(hipEventRecord(e0, ctx->stream));
                for (int r = 0; r < 2; ++r)
                    run_band_queries(sc, choices, spec, algo, kind, RFA_CACHED, cutoffs_u, d_qcut, dout, rest, &f);
                HIP_CHECK(hipEventRecord(e1, ctx->stream));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                return ms;
            };
            // words per lane (so lanes per pair and pairs per wave) for each bucket on its own
            uint8_t wpl[6] = {0, 0, 0, 0, 0, 0};
            bool any = false;
            for (uint32_t b = 0; b < 6; ++b) {
                const float t1 = time_runs(BandForce{b, 1, BAND_THREADS});
                if (t1 < 0.0f) continue;
                any = true;
                const float t2 = b ? time_runs(BandForce{b, 2, BAND_THREADS}) : -1.0f;
                wpl[b] = (t2 >= 0.0f && t2 < t1) ? 2 : 1;
            }
            if (any) {
                for (uint32_t b = 0; b < 6; ++b)
                    if (wpl[b]) ctx->band_wpl[b] = wpl[b];
                // then the workgroup size (pairs per workgroup) of the whole launch
                static const uint32_t threads[] = {64, 128, 256};
                float best_ms = -1.0f;
                uint32_t best_t = BAND_THREADS;
                for (uint32_t t : threads) {
                    const float ms = time_runs(BandForce{6, 0, t});
                    if (ms >= 0.0f && (best_ms < 0.0f || ms < best_ms)) {
                        best_ms = ms;
                        best_t = t;
                    }
                }
                ctx->band_threads = (uint16_t)best_t;
                if (chosen)
                    for (uint32_t b = 0; b < 6; ++b)
                        if (wpl[b]) {
                            chosen[2 * b] = wpl[b];
                            chosen[2 * b + 1] = best_t;
                        }
            }
        }
        catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

rfa_status rfa_topk(rfa_scorer* sc, const rfa_strings* choices, int kind, double cutoff_d, uint32_t k,
                    int64_t* out_index, double* out_score)
{
    if (!sc || !choices) return fail(RFA_ERR_INVALID_ARGUMENT, "scorer or choices is NULL");
    if (k && (!out_index || !out_score) && sc->q->count) return fail(RFA_ERR_INVALID_ARGUMENT, "outputs are NULL");
    return guarded([&] {
        std::lock_guard<std::mutex> lk(sc->ctx->mu);
        HIP_CHECK(hipSetDevice(sc->ctx->device));
        begin_call(sc->ctx);
        topk_impl(sc, choices, kind, cutoff_d, k, out_index, out_score);
    });
}

} // extern "C"
