// rfa_topk.cpp -- top-k extraction (rfa_topk) and the continuous-batching
// request server (rfa_server_*).
//
// Both run the README's extract() loop (README.md:153-214) on the device and
// keep the k best results by (score descending, choice index ascending):
//
//   fused path  queries of <= 256 one-byte units, metrics whose raw value is
//               an LCS length or a uniform-weight Levenshtein distance, a
//               similarity kind and a positive cutoff.  Choices are scored
//               from a length-sorted, lane-interleaved index of the choice
//               set (k_serve), candidates >= cutoff are kept per query and
//               ordered on the device (k_select).
//   dense path  everything else: rfa_cdist into a device matrix, chunked over
//               the choices, each chunk merged into the running top-k.  The
//               fused path also falls back to it for a query whose
//               candidates overflow their buffer.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "rfa_host.hpp"

namespace rfa {

ServeIndex::~ServeIndex()
{
    for (void* p : {(void*)d_data, (void*)d_base, (void*)d_dwords, (void*)d_glen_min, (void*)d_slot_len,
                    (void*)d_slot_idx})
        if (p) (void)hipFree(p);
}

namespace {

template <typename T>
T* to_device(const std::vector<T>& v)
{
    T* d = nullptr;
    const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    if (hipMalloc(&d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw ApiError(RFA_ERR_OUT_OF_MEMORY, "device allocation for the top-k index failed");
    }
    if (!v.empty()) HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// Sort by length, cut into groups of 64 lanes, interleave characters by dword.
std::shared_ptr<ServeIndex> build_index(const rfa_strings* ch)
{
    auto ix = std::make_shared<ServeIndex>();
    const uint64_t n = ch->count;
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ch->len(a) < ch->len(b); });
    const uint32_t ng = (uint32_t)((n + 63) / 64);
    ix->ngroups = ng;
    ix->glen_min.assign(ng, 0);
    ix->glen_max.assign(ng, 0);
    std::vector<uint64_t> base(ng);
    std::vector<uint32_t> dwords(ng), slot_len((size_t)ng * 64, 0), slot_idx((size_t)ng * 64, 0xffffffffu);
    uint64_t total = 0;
    for (uint32_t g = 0; g < ng; ++g) {
        uint64_t lo = ~0ull, hi = 0;
        for (uint32_t l = 0; l < 64; ++l) {
            const uint64_t k = (uint64_t)g * 64 + l;
            if (k >= n) break;
            const uint64_t len = ch->len(order[k]);
            lo = std::min(lo, len);
            hi = std::max(hi, len);
            slot_len[k] = (uint32_t)len;
            slot_idx[k] = order[k];
        }
        ix->glen_min[g] = (uint32_t)lo;
        ix->glen_max[g] = (uint32_t)hi;
        dwords[g] = (uint32_t)((hi + 3) / 4);
        base[g] = total;
        total += (uint64_t)dwords[g] * 64;
    }
    std::vector<uint32_t> data(total, 0u);
    uint8_t* bytes = (uint8_t*)data.data();
    for (uint64_t k = 0; k < n; ++k) {
        const uint32_t g = (uint32_t)(k / 64), l = (uint32_t)(k % 64);
        const uint64_t o = ch->h_off[order[k]];
        const uint32_t len = slot_len[k];
        for (uint32_t j = 0; j < len; ++j) bytes[(base[g] + (uint64_t)(j >> 2) * 64 + l) * 4 + (j & 3)] = ch->h_chars[o + j];
    }
    ix->data_bytes = total * 4;
    ix->d_data = to_device(data);
    ix->d_base = to_device(base);
    ix->d_dwords = to_device(dwords);
    ix->d_glen_min = to_device(ix->glen_min);
    ix->d_slot_len = to_device(slot_len);
    ix->d_slot_idx = to_device(slot_idx);
    return ix;
}

} // namespace

std::shared_ptr<ServeIndex> get_index(const rfa_strings* ch)
{
    auto* mut = const_cast<rfa_strings*>(ch); // a cache on an immutable set
    if (!mut->index) mut->index = build_index(ch);
    return mut->index;
}

// ---- fused path -------------------------------------------------------------

// Similarity on the normalized scale that the metric's best case (raw value
// from the length difference alone) reaches, as a bound type:
//   SUM  Indel / ratio / Levenshtein with Indel-like weights: 2 min / (m + n)
//   MAX  LCSseq / uniform Levenshtein:                           min / max
enum class Bound { NONE, SUM, MAX };

struct FusedKind {
    bool ok = false;
    int algo = 0;
    Bound bound = Bound::NONE;
    double norm_cutoff = 0; // cutoff on the [0, 1] scale
};

FusedKind fused_kind(int metric, const rfa_weights& w, int kind, double cutoff, int width)
{
    FusedKind f;
    if (width != 1 || !(cutoff > 0.0)) return f;
    const int algo = algo_for(metric, w);
    if (algo != ALGO_LCS && algo != ALGO_LEV) return f;
    if (metric == RFA_RATIO) {
        if (kind != RFA_SIMILARITY) return f;
        f.norm_cutoff = cutoff / 100.0;
    }
    else {
        if (kind != RFA_NORMALIZED_SIMILARITY) return f;
        f.norm_cutoff = cutoff;
    }
    f.ok = true;
    f.algo = algo;
    if (metric == RFA_LEVENSHTEIN && w.insert_cost == 0)
        f.bound = Bound::NONE; // every pair scores 1.0
    else if (metric == RFA_LCSSEQ || (metric == RFA_LEVENSHTEIN && algo == ALGO_LEV))
        f.bound = Bound::MAX;
    else
        f.bound = Bound::SUM;
    return f;
}

namespace {

// Lengths n that can reach the cutoff against a query of m units, widened
// (cutoff lowered by 1e-3 and one length either side) so rounding in the
// reference's cutoff chain cannot exclude a qualifying pair.
void length_hull(const FusedKind& f, uint64_t m, uint64_t maxlen, uint64_t* lo, uint64_t* hi)
{
    *lo = 0;
    *hi = maxlen;
    const double c = f.norm_cutoff - 1e-3;
    if (f.bound == Bound::NONE || m == 0 || !(c > 0.0)) return;
    double flo, fhi;
    if (f.bound == Bound::SUM) {
        flo = c * (double)m / (2.0 - c);
        fhi = (double)m * (2.0 - c) / c;
    }
    else {
        flo = c * (double)m;
        fhi = (double)m / c;
    }
    const double l = std::floor(flo) - 1.0, h = std::ceil(fhi) + 1.0;
    *lo = l <= 0.0 ? 0 : (uint64_t)l;
    *hi = h >= (double)maxlen ? maxlen : (uint64_t)h;
}

} // namespace

// Groups whose length span meets [lo, hi] (groups are sorted by length).
static void group_range(const ServeIndex& ix, uint64_t lo, uint64_t hi, uint32_t* glo, uint32_t* ghi)
{
    const auto b = std::lower_bound(ix.glen_max.begin(), ix.glen_max.end(), lo);
    const auto e = std::upper_bound(ix.glen_min.begin(), ix.glen_min.end(), hi);
    *glo = (uint32_t)(b - ix.glen_max.begin());
    *ghi = std::max(*glo, (uint32_t)(e - ix.glen_min.begin()));
}

// A batch of queries sorted by length, laid out for one host-to-device copy.
struct FusedPlan {
    struct Launch {
        uint32_t wq, q_begin, q_end, sub_first, nsub, gx;
    };
    std::vector<uint8_t> blob;
    size_t off_qoff = 0, off_qglo = 0, off_qghi = 0, off_sglo = 0, off_sghi = 0;
    std::vector<Launch> launches;
    uint32_t nq = 0;
};

static size_t align8(size_t x) { return (x + 7) & ~size_t(7); }

// queries[i] = (pointer, length), already sorted by length; each <= 256 units.
static void plan_fused(const ServeIndex& ix, const FusedKind& f, const std::vector<std::pair<const uint8_t*, uint32_t>>& qs,
                       uint64_t maxlen, int num_cus, FusedPlan& p)
{
    const uint32_t nq = (uint32_t)qs.size();
    p.nq = nq;
    p.launches.clear();
    uint64_t nchars = 0;
    for (auto& q : qs) nchars += q.second;
    std::vector<uint32_t> glo(nq), ghi(nq);
    for (uint32_t i = 0; i < nq; ++i) {
        uint64_t lo, hi;
        length_hull(f, qs[i].second, maxlen, &lo, &hi);
        group_range(ix, lo, hi, &glo[i], &ghi[i]);
    }
    // launches: one per word class (1, 2, 3-4 words), subsets of QPB queries
    std::vector<uint32_t> sglo, sghi;
    uint32_t i = 0;
    while (i < nq) {
        const uint32_t w = std::max<uint32_t>(1, (qs[i].second + 63) / 64);
        const uint32_t wq = w <= 1 ? 1 : (w <= 2 ? 2 : 4);
        uint32_t j = i;
        while (j < nq) {
            const uint32_t wj = std::max<uint32_t>(1, (qs[j].second + 63) / 64);
            if ((wj <= 1 ? 1 : (wj <= 2 ? 2 : 4)) != wq) break;
            ++j;
        }
        const uint32_t qpb = serve_queries_per_block(wq);
        FusedPlan::Launch L{wq, i, j, (uint32_t)sglo.size(), (j - i + qpb - 1) / qpb, 1};
        uint32_t R = 0;
        for (uint32_t s = 0; s < L.nsub; ++s) {
            uint32_t lo = ~0u, hi = 0;
            for (uint32_t q = i + s * qpb; q < std::min(j, i + (s + 1) * qpb); ++q) {
                if (glo[q] < ghi[q]) {
                    lo = std::min(lo, glo[q]);
                    hi = std::max(hi, ghi[q]);
                }
            }
            if (lo > hi) lo = hi = 0;
            sglo.push_back(lo);
            sghi.push_back(hi);
            R = std::max(R, hi - lo);
        }
        if (R) {
            const uint32_t target_waves = (uint32_t)num_cus * 8;
            const uint32_t want = std::max<uint32_t>(1, (target_waves + 4 * L.nsub - 1) / (4 * L.nsub));
            L.gx = std::min(want, (R + 3) / 4);
            p.launches.push_back(L);
        }
        i = j;
    }
    const uint32_t nsub = (uint32_t)sglo.size();
    p.off_qoff = align8(nchars);
    p.off_qglo = p.off_qoff + (size_t)(nq + 1) * 8;
    p.off_qghi = p.off_qglo + (size_t)nq * 4;
    p.off_sglo = p.off_qghi + (size_t)nq * 4;
    p.off_sghi = p.off_sglo + (size_t)nsub * 4;
    p.blob.assign(p.off_sghi + (size_t)nsub * 4, 0);
    uint64_t* qoff = (uint64_t*)(p.blob.data() + p.off_qoff);
    uint64_t pos = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        qoff[q] = pos;
        if (qs[q].second) std::memcpy(p.blob.data() + pos, qs[q].first, qs[q].second);
        pos += qs[q].second;
    }
    qoff[nq] = pos;
    if (nq) {
        std::memcpy(p.blob.data() + p.off_qglo, glo.data(), nq * 4);
        std::memcpy(p.blob.data() + p.off_qghi, ghi.data(), nq * 4);
    }
    if (nsub) {
        std::memcpy(p.blob.data() + p.off_sglo, sglo.data(), nsub * 4);
        std::memcpy(p.blob.data() + p.off_sghi, sghi.data(), nsub * 4);
    }
}

// Enqueue scoring + selection of a planned batch; d_blob already holds p.blob.
static void enqueue_fused(const ServeIndex& ix, const FusedPlan& p, const ScoreSpec& spec, int algo, double cutoff,
                          uint32_t k, const uint8_t* d_blob, uint32_t* d_counts, double* d_cand_s, int64_t* d_cand_i,
                          uint32_t cap, double* d_out_s, int64_t* d_out_i, hipStream_t s)
{
    if (!p.nq) return;
    HIP_CHECK(hipMemsetAsync(d_counts, 0, (size_t)p.nq * 4, s));
    ServeArgs a{};
    a.gdata = ix.d_data;
    a.gbase = ix.d_base;
    a.gdwords = ix.d_dwords;
    a.glen_min = ix.d_glen_min;
    a.slot_len = ix.d_slot_len;
    a.slot_idx = ix.d_slot_idx;
    a.qchars = d_blob;
    a.qoff = (const uint64_t*)(d_blob + p.off_qoff);
    a.qglo = (const uint32_t*)(d_blob + p.off_qglo);
    a.qghi = (const uint32_t*)(d_blob + p.off_qghi);
    a.spec = spec;
    a.cutoff = cutoff;
    a.counts = d_counts;
    a.cand_score = d_cand_s;
    a.cand_index = d_cand_i;
    a.cap = cap;
    for (const auto& L : p.launches) {
        a.q_begin = L.q_begin;
        a.q_end = L.q_end;
        a.sub_glo = (const uint32_t*)(d_blob + p.off_sglo) + L.sub_first;
        a.sub_ghi = (const uint32_t*)(d_blob + p.off_sghi) + L.sub_first;
        HIP_CHECK(launch_serve(algo, L.wq, a, L.gx, L.nsub, s));
    }
    if (k) {
        SelectArgs sa{};
        sa.lscore = d_cand_s;
        sa.lindex = d_cand_i;
        sa.lcount = d_counts;
        sa.lcap = cap;
        sa.lstride = cap;
        sa.cutoff = cutoff;
        sa.k = k;
        sa.out_score = d_out_s;
        sa.out_index = d_out_i;
        HIP_CHECK(launch_select(sa, p.nq, s));
    }
}

// ---- cdist of a few queries: one pair per lane --------------------------------

// The plan of a scorer's queries against one choice set, uploaded once (the
// queries and the set are immutable, so repeated calls reuse it).
struct FewQueryPlan {
    std::shared_ptr<ServeIndex> ix; // identifies the set (and keeps its index alive)
    FusedPlan p;
    size_t off_qrow = 0;
    uint8_t* d_blob = nullptr;
    ~FewQueryPlan()
    {
        if (d_blob) (void)hipFree(d_blob);
    }
};

bool cdist_few_queries(rfa_scorer* sc, const rfa_strings* ch, const ScoreSpec& spec, int algo, void* dout,
                       const uint64_t* d_qcut)
{
    rfa_context* ctx = sc->ctx;
    const uint64_t nq = sc->q->count, nc = ch->count;
    if (nq == 0 || nq > 16 || nc == 0 || nc >= 0xffffffffull) return false;
    if (ch->width != 1 || sc->q->width != 1 || sc->q->max_len > 256) return false;
    if (ch->view) return false; // chunk views of another set carry no host copy to index
    if (algo != ALGO_LCS && algo != ALGO_LEV) return false;
    auto ix = get_index(ch);
    auto* cached = static_cast<FewQueryPlan*>(sc->few_plan.get());
    if (!cached || cached->ix != ix) {
        auto plan = std::make_shared<FewQueryPlan>();
        plan->ix = ix;
        std::vector<uint32_t> order(nq);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return sc->q->len(a) < sc->q->len(b); });
        std::vector<std::pair<const uint8_t*, uint32_t>> qs(nq);
        for (uint64_t i = 0; i < nq; ++i)
            qs[i] = {sc->q->h_chars.data() + sc->q->h_off[order[i]], (uint32_t)sc->q->len(order[i])};
        FusedKind f;
        f.ok = true;
        f.algo = algo;
        f.bound = Bound::NONE; // every result is wanted: all groups
        plan_fused(*ix, f, qs, ch->max_len, ctx->num_cus, plan->p);
        plan->off_qrow = align8(plan->p.blob.size());
        plan->p.blob.resize(plan->off_qrow + nq * 4);
        std::memcpy(plan->p.blob.data() + plan->off_qrow, order.data(), nq * 4);
        plan->d_blob = (uint8_t*)to_device(plan->p.blob);
        sc->few_plan = plan;
        cached = plan.get();
    }
    const FusedPlan& p = cached->p;
    const size_t off_qrow = cached->off_qrow;
    hipStream_t s = ctx->stream;
    const uint8_t* d_blob = cached->d_blob;
    ServeArgs a{};
    a.gdata = ix->d_data;
    a.gbase = ix->d_base;
    a.gdwords = ix->d_dwords;
    a.glen_min = ix->d_glen_min;
    a.slot_len = ix->d_slot_len;
    a.slot_idx = ix->d_slot_idx;
    a.qchars = d_blob;
    a.qoff = (const uint64_t*)(d_blob + p.off_qoff);
    a.qglo = (const uint32_t*)(d_blob + p.off_qglo);
    a.qghi = (const uint32_t*)(d_blob + p.off_qghi);
    a.spec = spec;
    a.dense_out = (uint64_t*)dout;
    a.qrow = (const uint32_t*)(d_blob + off_qrow);
    a.dense_stride = nc;
    a.qcut = d_qcut;
    for (const auto& L : p.launches) {
        a.q_begin = L.q_begin;
        a.q_end = L.q_end;
        a.sub_glo = (const uint32_t*)(d_blob + p.off_sglo) + L.sub_first;
        a.sub_ghi = (const uint32_t*)(d_blob + p.off_sghi) + L.sub_first;
        HIP_CHECK(launch_serve(algo, L.wq, a, L.gx, L.nsub, s));
    }
    return true;
}

// ---- dense path -------------------------------------------------------------

// Top-k of every query of `sc` against `ch` through rfa_cdist, chunked over
// the choices; results to host memory.
static void topk_dense(rfa_scorer* sc, const rfa_strings* ch, int kind, double cutoff, uint32_t k, int64_t* out_index,
                       double* out_score)
{
    rfa_context* ctx = sc->ctx;
    const uint64_t nq = sc->q->count, nc = ch->count;
    if (!nq || !k) return;
    hipStream_t s = ctx->stream;
    constexpr uint64_t CHUNK_BYTES = 512ull << 20;
    const uint64_t C = std::max<uint64_t>(1, std::min<uint64_t>(nc, CHUNK_BYTES / 8 / nq));
    double* run_s[2];
    int64_t* run_i[2];
    for (int b = 0; b < 2; ++b) {
        uint8_t* base = (uint8_t*)ctx->topk_run[b].get((size_t)nq * k * 16);
        run_s[b] = (double*)base;
        run_i[b] = (int64_t*)(base + (size_t)nq * k * 8);
    }
    int cur = 0;
    bool have_prev = false;
    if (nc == 0) {
        std::fill(out_index, out_index + nq * k, (int64_t)-1);
        std::fill(out_score, out_score + nq * k, 0.0);
        return;
    }
    for (uint64_t c0 = 0; c0 < nc; c0 += C) {
        const uint64_t cn = std::min(C, nc - c0);
        rfa_strings view;
        view.ctx = ctx;
        view.width = ch->width;
        view.count = cn;
        view.d_chars = ch->d_chars;
        view.d_off = ch->d_off + c0;
        view.h_off.assign(ch->h_off.begin() + c0, ch->h_off.begin() + c0 + cn + 1);
        view.max_len = ch->max_len;
        view.view = true;
        double* dense = (double*)ctx->topk_scores.get((size_t)nq * cn * 8);
        cdist_impl(sc, &view, kind, RFA_CACHED, 0, cutoff, dense, 1);
        SelectArgs sa{};
        if (have_prev) {
            sa.lscore = run_s[cur];
            sa.lindex = run_i[cur];
            sa.lfixed = k;
            sa.lstride = k;
        }
        sa.dscore = dense;
        sa.dstride = cn;
        sa.dn = cn;
        sa.dbase = (int64_t)c0;
        sa.cutoff = cutoff;
        sa.k = k;
        sa.out_score = run_s[cur ^ 1];
        sa.out_index = run_i[cur ^ 1];
        HIP_CHECK(launch_select(sa, (uint32_t)nq, s));
        cur ^= 1;
        have_prev = true;
    }
    HIP_CHECK(hipMemcpyAsync(out_score, run_s[cur], (size_t)nq * k * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(out_index, run_i[cur], (size_t)nq * k * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// Top-k of the token-set ratios: their choice token tables are cached per string set,
// so the whole set is scored at once (no chunk views), then selected.
static void topk_tset(rfa_scorer* sc, const rfa_strings* ch, int kind, double cutoff, uint32_t k, int64_t* out_index,
                      double* out_score)
{
    rfa_context* ctx = sc->ctx;
    const uint64_t nq = sc->q->count, nc = ch->count;
    if (!nq || !k) return;
    if (nc == 0) {
        std::fill(out_index, out_index + nq * k, (int64_t)-1);
        std::fill(out_score, out_score + nq * k, 0.0);
        return;
    }
    constexpr uint64_t DENSE_MAX = 16ull << 30;
    if (nq * nc * 8 > DENSE_MAX)
        throw ApiError(RFA_ERR_UNSUPPORTED, "top-k of a token-set ratio or WRatio over more than 2^31 query-choice pairs");
    hipStream_t s = ctx->stream;
    double* dense = (double*)ctx->topk_scores.get((size_t)nq * nc * 8);
    if (sc->wratio)
        cdist_wratio(sc, ch, kind, RFA_CACHED, cutoff, dense, 1);
    else
        cdist_tset(sc, ch, kind, RFA_CACHED, cutoff, dense, 1);
    uint8_t* base = (uint8_t*)ctx->topk_run[0].get((size_t)nq * k * 16);
    SelectArgs sa{};
    sa.dscore = dense;
    sa.dstride = nc;
    sa.dn = nc;
    sa.dbase = 0;
    sa.cutoff = cutoff;
    sa.k = k;
    sa.out_score = (double*)base;
    sa.out_index = (int64_t*)(base + (size_t)nq * k * 8);
    HIP_CHECK(launch_select(sa, (uint32_t)nq, s));
    HIP_CHECK(hipMemcpyAsync(out_score, sa.out_score, (size_t)nq * k * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(out_index, sa.out_index, (size_t)nq * k * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// Top-k for a subset of a scorer's queries through the dense path.
static void topk_dense_subset(rfa_context* ctx, int metric, const rfa_weights& w, double pw,
                              const std::vector<std::pair<const uint8_t*, uint32_t>>& qs, int width,
                              const rfa_strings* ch, int kind, double cutoff, uint32_t k, int64_t* out_index,
                              double* out_score)
{
    std::vector<uint8_t> chars;
    std::vector<uint64_t> off{0};
    for (auto& q : qs) {
        chars.insert(chars.end(), q.first, q.first + (size_t)q.second * width);
        off.push_back(off.back() + q.second);
    }
    rfa_strings* tmp = make_strings(ctx, chars.data(), width, off.data(), qs.size());
    rfa_scorer* sc = nullptr;
    try {
        sc = make_scorer(ctx, metric, tmp, w, pw);
        topk_dense(sc, ch, kind, cutoff, k, out_index, out_score);
    }
    catch (...) {
        free_scorer(sc);
        free_strings(tmp);
        throw;
    }
    free_scorer(sc);
    free_strings(tmp);
}

static bool is_token_set(int metric)
{
    return metric == RFA_TOKEN_SET_RATIO || metric == RFA_TOKEN_RATIO || metric == RFA_PARTIAL_TOKEN_SET_RATIO ||
           metric == RFA_PARTIAL_TOKEN_RATIO;
}

static void validate_topk(int metric, int kind)
{
    validate_kind(metric, kind, RFA_CACHED);
    if (is_token_set(metric) && kind != RFA_SIMILARITY)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "token ratios only provide similarity");
    const bool sim_double = kind == RFA_NORMALIZED_SIMILARITY ||
                            (kind == RFA_SIMILARITY &&
                             (metric == RFA_JARO || metric == RFA_JARO_WINKLER || metric == RFA_RATIO ||
                              metric == RFA_PARTIAL_RATIO || is_token_set(metric)));
    if (!sim_double)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT,
                       "top-k ranks double similarities: use a normalized similarity, or the similarity of "
                       "Jaro, Jaro-Winkler, ratio or partial_ratio");
}

void topk_impl(rfa_scorer* sc, const rfa_strings* ch, int kind, double cutoff, uint32_t k, int64_t* out_index,
               double* out_score)
{
    rfa_context* ctx = sc->ctx;
    validate_topk(sc->metric, kind);
    if (ch->width != sc->q->width)
        throw ApiError(RFA_ERR_INVALID_ARGUMENT, "queries and choices must have the same char width");
    if (sc->token_set || sc->wratio) {
        topk_tset(sc, ch, kind, cutoff, k, out_index, out_score);
        return;
    }
    ch = scorer_choices(sc, ch);
    const uint64_t nq = sc->q->count, nc = ch->count;
    if (!nq || !k) return;
    if (nc >= 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 - 1 choices");
    // (QRatio's empty-string rule is applied by the dense path's epilogue only)
    const FusedKind f = sc->empty_zero ? FusedKind{} : fused_kind(sc->metric, sc->w, kind, cutoff, sc->q->width);
    std::vector<uint32_t> fused, dense;
    for (uint64_t q = 0; q < nq; ++q) {
        if (f.ok && sc->q->len(q) <= 256 && nc)
            fused.push_back((uint32_t)q);
        else
            dense.push_back((uint32_t)q);
    }
    if (!fused.empty()) {
        auto ix = get_index(ch);
        const int algo = f.algo;
        const ScoreSpec spec = [&] {
            ScoreSpec s{};
            s.metric = sc->metric;
            s.kind = kind;
            s.semantics = RFA_CACHED;
            s.algo = algo;
            s.cutoff_d = cutoff;
            s.ins = sc->w.insert_cost;
            s.del = sc->w.delete_cost;
            s.rep = sc->w.replace_cost;
            s.prefix_weight = sc->prefix_weight;
            s.band_k = ~0ull;
            return s;
        }();
        std::stable_sort(fused.begin(), fused.end(), [&](uint32_t a, uint32_t b) { return sc->q->len(a) < sc->q->len(b); });
        hipStream_t s = ctx->stream;
        // query chunks bound the candidate memory and the grid's y extent
        constexpr uint32_t QCHUNK = 4096;
        const uint32_t cap = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nc, 1), 16384);
        std::vector<double> hs;
        std::vector<int64_t> hi;
        std::vector<uint32_t> hc;
        for (size_t c0 = 0; c0 < fused.size(); c0 += QCHUNK) {
            const size_t cn = std::min<size_t>(QCHUNK, fused.size() - c0);
            std::vector<std::pair<const uint8_t*, uint32_t>> qs(cn);
            for (size_t i = 0; i < cn; ++i) {
                const uint32_t q = fused[c0 + i];
                qs[i] = {sc->q->h_chars.data() + sc->q->h_off[q], (uint32_t)sc->q->len(q)};
            }
            FusedPlan p;
            plan_fused(*ix, f, qs, ch->max_len, ctx->num_cus, p);
            uint8_t* d_blob = (uint8_t*)ctx->topk_q.get(p.blob.size());
            HIP_CHECK(hipMemcpyAsync(d_blob, p.blob.data(), p.blob.size(), hipMemcpyHostToDevice, s));
            uint32_t* d_counts = (uint32_t*)ctx->topk_counts.get(cn * 4);
            double* d_cs = (double*)ctx->topk_cand_s.get(cn * cap * 8);
            int64_t* d_ci = (int64_t*)ctx->topk_cand_i.get(cn * cap * 8);
            uint8_t* d_out = (uint8_t*)ctx->topk_out.get(cn * k * 16);
            double* d_os = (double*)d_out;
            int64_t* d_oi = (int64_t*)(d_out + cn * k * 8);
            enqueue_fused(*ix, p, spec, algo, cutoff, k, d_blob, d_counts, d_cs, d_ci, cap, d_os, d_oi, s);
            hs.resize(cn * k);
            hi.resize(cn * k);
            hc.resize(cn);
            HIP_CHECK(hipMemcpyAsync(hs.data(), d_os, cn * k * 8, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(hi.data(), d_oi, cn * k * 8, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipMemcpyAsync(hc.data(), d_counts, cn * 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            for (size_t i = 0; i < cn; ++i) {
                const uint32_t q = fused[c0 + i];
                if (hc[i] > cap) {
                    dense.push_back(q); // candidates overflowed: redo densely
                    continue;
                }
                std::memcpy(out_score + (size_t)q * k, hs.data() + i * k, k * 8);
                std::memcpy(out_index + (size_t)q * k, hi.data() + i * k, k * 8);
            }
        }
    }
    if (!dense.empty()) {
        std::sort(dense.begin(), dense.end());
        if (dense.size() == nq) {
            topk_dense(sc, ch, kind, cutoff, k, out_index, out_score);
        }
        else {
            std::vector<std::pair<const uint8_t*, uint32_t>> qs;
            for (uint32_t q : dense)
                qs.push_back({sc->q->h_chars.data() + sc->q->h_off[q] * sc->q->width, (uint32_t)sc->q->len(q)});
            std::vector<int64_t> ri(dense.size() * k);
            std::vector<double> rs(dense.size() * k);
            topk_dense_subset(ctx, sc->metric, sc->w, sc->prefix_weight, qs, sc->q->width, ch, kind, cutoff, k,
                              ri.data(), rs.data());
            for (size_t i = 0; i < dense.size(); ++i) {
                std::memcpy(out_score + (size_t)dense[i] * k, rs.data() + i * k, k * 8);
                std::memcpy(out_index + (size_t)dense[i] * k, ri.data() + i * k, k * 8);
            }
        }
    }
}

} // namespace rfa

// ---- request server ---------------------------------------------------------------

using namespace rfa;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    std::vector<uint8_t> q; // code units, server width
    uint32_t len = 0;
    bool done = false;
    rfa_status st = RFA_OK;
    std::string err;
    std::vector<int64_t> idx;
    std::vector<double> score;
    Clock::time_point t_done;
};

struct Slot {
    hipStream_t stream = nullptr;
    DevBuf blob, counts, cand_s, cand_i, out;
    uint8_t* h_out = nullptr; // pinned: scores, indices, counts
    std::thread th;
};

} // namespace

struct rfa_server {
    rfa_context* ctx = nullptr;
    rfa_server_config cfg{};
    rfa_weights w{1, 1, 1};
    rfa_strings* choices = nullptr; // own copy
    std::shared_ptr<ServeIndex> ix;
    FusedKind f;
    ScoreSpec spec{};
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    rfa_server_stats stats{};
};

namespace {

bool fused_query(const rfa_server* srv, const Request& r) { return srv->f.ok && r.len <= 256; }

// Runs one admitted batch; fills r->idx/score (or st/err) of every request.
void run_batch(rfa_server* srv, Slot& slot, const std::vector<Request*>& batch)
{
    const uint32_t k = srv->cfg.k;
    const int width = srv->choices->width;
    const bool fused = fused_query(srv, *batch[0]);
    std::vector<std::pair<const uint8_t*, uint32_t>> qs(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) qs[i] = {batch[i]->q.data(), batch[i]->len};
    for (Request* r : batch) {
        r->idx.assign(k, -1);
        r->score.assign(k, 0.0);
    }
    std::vector<uint32_t> redo;
    if (fused && srv->choices->count) {
        // sorted by length: subsets of one workgroup share word counts and length ranges
        std::vector<uint32_t> ord(batch.size());
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return qs[a].second < qs[b].second; });
        std::vector<std::pair<const uint8_t*, uint32_t>> sq(batch.size());
        for (size_t i = 0; i < ord.size(); ++i) sq[i] = qs[ord[i]];
        FusedPlan p;
        plan_fused(*srv->ix, srv->f, sq, srv->choices->max_len, srv->ctx->num_cus, p);
        const uint32_t nq = p.nq, cap = srv->cfg.candidate_cap;
        hipStream_t s = slot.stream;
        uint8_t* d_blob = (uint8_t*)slot.blob.get(p.blob.size());
        HIP_CHECK(hipMemcpyAsync(d_blob, p.blob.data(), p.blob.size(), hipMemcpyHostToDevice, s));
        uint32_t* d_counts = (uint32_t*)slot.counts.get((size_t)nq * 4);
        double* d_cs = (double*)slot.cand_s.get((size_t)nq * cap * 8);
        int64_t* d_ci = (int64_t*)slot.cand_i.get((size_t)nq * cap * 8);
        uint8_t* d_out = (uint8_t*)slot.out.get((size_t)nq * k * 16 + (size_t)nq * 4);
        double* d_os = (double*)d_out;
        int64_t* d_oi = (int64_t*)(d_out + (size_t)nq * k * 8);
        enqueue_fused(*srv->ix, p, srv->spec, srv->f.algo, srv->cfg.score_cutoff, k, d_blob, d_counts, d_cs, d_ci,
                      cap, d_os, d_oi, s);
        HIP_CHECK(hipMemcpyAsync(d_out + (size_t)nq * k * 16, d_counts, (size_t)nq * 4, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipMemcpyAsync(slot.h_out, d_out, (size_t)nq * k * 16 + (size_t)nq * 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        const double* hs = (const double*)slot.h_out;
        const int64_t* hi = (const int64_t*)(slot.h_out + (size_t)nq * k * 8);
        const uint32_t* hc = (const uint32_t*)(slot.h_out + (size_t)nq * k * 16);
        for (uint32_t i = 0; i < nq; ++i) {
            const uint32_t b = ord[i];
            if (hc[i] > cap) {
                redo.push_back(b);
                continue;
            }
            std::memcpy(batch[b]->score.data(), hs + (size_t)i * k, k * 8);
            std::memcpy(batch[b]->idx.data(), hi + (size_t)i * k, k * 8);
        }
    }
    else if (srv->choices->count) {
        for (uint32_t b = 0; b < batch.size(); ++b) redo.push_back(b);
    }
    if (!redo.empty()) {
        std::vector<std::pair<const uint8_t*, uint32_t>> rq;
        for (uint32_t b : redo) rq.push_back(qs[b]);
        std::vector<int64_t> ri(redo.size() * k);
        std::vector<double> rs(redo.size() * k);
        {
            std::lock_guard<std::mutex> lk(srv->ctx->mu);
            begin_call(srv->ctx);
            topk_dense_subset(srv->ctx, srv->cfg.metric, srv->w, srv->cfg.prefix_weight, rq, width, srv->choices,
                              srv->cfg.kind, srv->cfg.score_cutoff, k, ri.data(), rs.data());
        }
        for (size_t i = 0; i < redo.size(); ++i) {
            std::memcpy(batch[redo[i]]->score.data(), rs.data() + i * k, k * 8);
            std::memcpy(batch[redo[i]]->idx.data(), ri.data() + i * k, k * 8);
        }
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stats.dense_fallbacks += redo.size();
    }
}

void slot_main(rfa_server* srv, Slot* slot)
{
    (void)hipSetDevice(srv->ctx->device);
    for (;;) {
        std::vector<Request*> batch;
        std::vector<uint64_t> tickets;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] { return srv->stop || !srv->pending.empty(); });
            if (srv->pending.empty()) return; // stop with nothing left
            // admission: arrival order, up to max_batch requests (bounded LDS per workgroup
            // and candidate memory per slot); a batch holds either fused-path requests only
            // or dense-path requests only, by the kind of the oldest pending request
            Request& first = srv->reqs.at(srv->pending.front());
            const bool fused = fused_query(srv, first);
            const size_t limit = srv->cfg.max_batch;
            for (auto it = srv->pending.begin(); it != srv->pending.end() && batch.size() < limit;) {
                Request& r = srv->reqs.at(*it);
                if (fused_query(srv, r) == fused) {
                    batch.push_back(&r);
                    tickets.push_back(*it);
                    it = srv->pending.erase(it);
                }
                else {
                    ++it;
                }
            }
            srv->stats.batches += 1;
            srv->stats.requests += batch.size();
            srv->stats.max_batch_seen = std::max<uint64_t>(srv->stats.max_batch_seen, batch.size());
        }
        rfa_status st = RFA_OK;
        std::string err;
        try {
            run_batch(srv, *slot, batch);
        }
        catch (const ApiError& e) {
            st = e.st;
            err = e.what();
        }
        catch (const HipError& e) {
            st = RFA_ERR_HIP;
            err = e.what();
        }
        catch (const std::exception& e) {
            st = RFA_ERR_INVALID_ARGUMENT;
            err = e.what();
        }
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (Request* r : batch) {
                r->done = true;
                r->st = st;
                r->err = err;
                r->t_done = now;
            }
        }
        srv->cv_done.notify_all();
    }
}

void destroy_server(rfa_server* srv)
{
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto& s : srv->slots)
        if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(srv->ctx->device);
    for (auto& s : srv->slots) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        for (DevBuf* b : {&s->blob, &s->counts, &s->cand_s, &s->cand_i, &s->out}) b->release();
        if (s->h_out) (void)hipHostFree(s->h_out);
        if (s->stream) (void)hipStreamDestroy(s->stream);
    }
    srv->ix.reset();
    {
        std::lock_guard<std::mutex> lk(srv->ctx->mu);
        (void)hipStreamSynchronize(srv->ctx->stream);
        free_strings(srv->choices);
    }
    delete srv;
}

rfa_status wait_ticket(rfa_server* srv, uint64_t ticket, int64_t* out_index, double* out_score, Clock::time_point* t_done)
{
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return fail(RFA_ERR_INVALID_ARGUMENT, "unknown ticket");
    Request& r = it->second; // element references survive rehashing, iterators do not
    srv->cv_done.wait(lk, [&] { return r.done; });
    const rfa_status st = r.st;
    if (st == RFA_OK) {
        if (out_index) std::memcpy(out_index, r.idx.data(), r.idx.size() * 8);
        if (out_score) std::memcpy(out_score, r.score.data(), r.score.size() * 8);
        if (t_done) *t_done = r.t_done;
    }
    std::string err = r.err;
    srv->reqs.erase(ticket);
    return st == RFA_OK ? RFA_OK : fail(st, err);
}

uint64_t submit(rfa_server* srv, const void* query, uint64_t len)
{
    const int width = srv->choices->width;
    Request r;
    r.len = (uint32_t)len;
    r.q.assign((const uint8_t*)query, (const uint8_t*)query + len * width);
    uint64_t t;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        t = srv->next_ticket++;
        srv->reqs.emplace(t, std::move(r));
        srv->pending.push_back(t);
    }
    srv->cv_work.notify_one();
    return t;
}

} // namespace

extern "C" {

rfa_status rfa_server_create(rfa_context* ctx, const rfa_strings* choices, const rfa_server_config* cfg,
                             rfa_server** out)
{
    if (!ctx || !choices || !cfg || !out) return fail(RFA_ERR_INVALID_ARGUMENT, "ctx, choices, cfg or out is NULL");
    return guarded([&] {
        if (cfg->metric < RFA_LEVENSHTEIN || cfg->metric > RFA_PARTIAL_RATIO)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "unknown metric");
        validate_topk(cfg->metric, cfg->kind);
        if (cfg->k == 0) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "k must be positive");
        if (cfg->max_batch == 0 || cfg->max_batch > 4096)
            throw ApiError(RFA_ERR_INVALID_ARGUMENT, "max_batch must be in 1..4096");
        if (cfg->slots == 0 || cfg->slots > 8) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "slots must be in 1..8");
        if (cfg->candidate_cap == 0) throw ApiError(RFA_ERR_INVALID_ARGUMENT, "candidate_cap must be positive");
        if (choices->count >= 0xffffffffull) throw ApiError(RFA_ERR_UNSUPPORTED, "more than 2^32 - 1 choices");
        rfa_weights w = cfg->metric == RFA_LEVENSHTEIN ? cfg->weights : rfa_weights{1, 1, 1};
        if (cfg->metric == RFA_LEVENSHTEIN && w.insert_cost == 0 && w.delete_cost == 0 && w.replace_cost == 0)
            w = rfa_weights{1, 1, 1}; // zero-initialized config: the reference default
        if (algo_for(cfg->metric, w) < 0 || algo_for(cfg->metric, w) == ALGO_WLEV)
            throw ApiError(RFA_ERR_UNSUPPORTED, "the server takes Levenshtein weights {w,w,w} or Indel-like "
                                                "{w,w,>=2w} (score other weights with rfa_topk or rfa_cdist)");
        auto* srv = new rfa_server();
        srv->ctx = ctx;
        srv->cfg = *cfg;
        srv->w = w;
        try {
            {
                std::lock_guard<std::mutex> lk(ctx->mu);
                HIP_CHECK(hipSetDevice(ctx->device));
                srv->choices = make_strings(ctx, choices->h_chars.data(), choices->width, choices->h_off.data(),
                                            choices->count);
            }
            srv->f = fused_kind(cfg->metric, w, cfg->kind, cfg->score_cutoff, choices->width);
            if (srv->f.ok) srv->ix = get_index(srv->choices);
            ScoreSpec& s = srv->spec;
            s.metric = cfg->metric;
            s.kind = cfg->kind;
            s.semantics = RFA_CACHED;
            s.algo = srv->f.algo;
            s.cutoff_d = cfg->score_cutoff;
            s.ins = w.insert_cost;
            s.del = w.delete_cost;
            s.rep = w.replace_cost;
            s.prefix_weight = cfg->prefix_weight;
            s.band_k = ~0ull;
            for (uint32_t i = 0; i < cfg->slots; ++i) {
                auto sl = std::make_unique<Slot>();
                HIP_CHECK(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
                const size_t nq = cfg->max_batch;
                HIP_CHECK(hipHostMalloc((void**)&sl->h_out, nq * cfg->k * 16 + nq * 4, hipHostMallocDefault));
                if (srv->f.ok) { // reserve the batch's device memory up front
                    sl->blob.get(nq * 256 + (nq + 1) * 8 + nq * 16 + 64);
                    sl->counts.get(nq * 4);
                    sl->cand_s.get(nq * cfg->candidate_cap * 8);
                    sl->cand_i.get(nq * cfg->candidate_cap * 8);
                    sl->out.get(nq * cfg->k * 16 + nq * 4);
                }
                srv->slots.push_back(std::move(sl));
            }
            for (auto& sl : srv->slots) sl->th = std::thread(slot_main, srv, sl.get());
        }
        catch (...) {
            destroy_server(srv);
            throw;
        }
        *out = srv;
    });
}

rfa_status rfa_server_destroy(rfa_server* srv)
{
    if (!srv) return RFA_OK;
    return guarded([&] { destroy_server(srv); });
}

rfa_status rfa_server_submit(rfa_server* srv, const void* query, uint64_t len, uint64_t* ticket)
{
    if (!srv || !ticket) return fail(RFA_ERR_INVALID_ARGUMENT, "srv or ticket is NULL");
    if (len && !query) return fail(RFA_ERR_INVALID_ARGUMENT, "query is NULL");
    return guarded([&] { *ticket = submit(srv, query, len); });
}

rfa_status rfa_server_wait(rfa_server* srv, uint64_t ticket, int64_t* out_index, double* out_score)
{
    if (!srv) return fail(RFA_ERR_INVALID_ARGUMENT, "srv is NULL");
    return wait_ticket(srv, ticket, out_index, out_score, nullptr);
}

rfa_status rfa_server_replay(rfa_server* srv, const void* chars, const uint64_t* offsets, uint64_t count,
                             const double* arrival_s, double* latency_s, int64_t* out_index, double* out_score)
{
    if (!srv || !offsets || !arrival_s || !latency_s) return fail(RFA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count && offsets[count] && !chars) return fail(RFA_ERR_INVALID_ARGUMENT, "chars is NULL");
    for (uint64_t i = 0; i < count; ++i)
        if (offsets[i + 1] < offsets[i]) return fail(RFA_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    const int width = srv->choices->width;
    const uint32_t k = srv->cfg.k;
    std::vector<uint64_t> tickets(count);
    std::vector<Clock::time_point> due(count);
    const auto t0 = Clock::now();
    rfa_status st = RFA_OK;
    std::string err;
    // the feeder submits at the scheduled times; this thread collects in order
    std::thread feeder([&] {
        for (uint64_t i = 0; i < count; ++i) {
            due[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
            std::this_thread::sleep_until(due[i]);
            const uint64_t t = submit(srv, (const uint8_t*)chars + offsets[i] * width, offsets[i + 1] - offsets[i]);
            std::lock_guard<std::mutex> lk(srv->mu);
            tickets[i] = t;
            srv->cv_done.notify_all();
        }
    });
    for (uint64_t i = 0; i < count; ++i) {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_done.wait(lk, [&] { return tickets[i] != 0; });
        }
        Clock::time_point td;
        const rfa_status s = wait_ticket(srv, tickets[i], out_index ? out_index + i * k : nullptr,
                                         out_score ? out_score + i * k : nullptr, &td);
        if (s != RFA_OK && st == RFA_OK) {
            st = s;
            err = g_err;
        }
        latency_s[i] = s == RFA_OK ? std::chrono::duration<double>(td - due[i]).count() : -1.0;
    }
    feeder.join();
    return st == RFA_OK ? RFA_OK : fail(st, err);
}

rfa_status rfa_server_get_stats(rfa_server* srv, rfa_server_stats* out)
{
    if (!srv || !out) return fail(RFA_ERR_INVALID_ARGUMENT, "srv or out is NULL");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    return RFA_OK;
}

} // extern "C"
