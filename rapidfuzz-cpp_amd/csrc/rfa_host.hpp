// rfa_host.hpp -- host-side state shared by the C ABI translation units:
// error plumbing, growable device buffers, and the definitions of the opaque
// handle types declared in rapidfuzz_amd.h.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "rfa_internal.hpp"

namespace rfa {

extern thread_local std::string g_err;

inline rfa_status fail(rfa_status st, const std::string& msg)
{
    g_err = msg;
    return st;
}

struct HipError : std::runtime_error {
    hipError_t e;
    HipError(hipError_t e_, const char* what) : std::runtime_error(std::string(what) + ": " + hipGetErrorString(e_)), e(e_) {}
};

#define HIP_CHECK(x)                                                                                               \
    do {                                                                                                           \
        hipError_t e_ = (x);                                                                                       \
        if (e_ != hipSuccess) throw ::rfa::HipError(e_, #x);                                                       \
    } while (0)

struct ApiError : std::runtime_error {
    rfa_status st;
    ApiError(rfa_status s, const std::string& m) : std::runtime_error(m), st(s) {}
};

// Device buffer that grows on demand (kept by its owner between calls).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes)
    {
        if (bytes > cap) {
            if (p) HIP_CHECK(hipFree(p));
            p = nullptr;
            cap = 0;
            size_t want = std::max<size_t>(bytes, 256);
            if (hipMalloc(&p, want) != hipSuccess) {
                (void)hipGetLastError();
                throw ApiError(RFA_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(want) + " bytes failed");
            }
            cap = want;
        }
        return p;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Host staging of asynchronous uploads: the source vectors stay alive until
// the stream has passed them (see begin_call).
struct Keep {
    std::vector<std::shared_ptr<void>> items;
};

template <typename T>
T* upload(Keep& keep, DevBuf& b, const std::vector<T>& v, hipStream_t s)
{
    T* d = (T*)b.get(std::max<size_t>(v.size(), 1) * sizeof(T));
    if (!v.empty()) {
        auto copy = std::make_shared<std::vector<T>>(v);
        HIP_CHECK(hipMemcpyAsync(d, copy->data(), copy->size() * sizeof(T), hipMemcpyHostToDevice, s));
        keep.items.push_back(copy);
    }
    return d;
}

// Pinned host staging for per-call uploads.  Two slots are used in turn, and a
// slot is rewritten only after the copy that last read it has completed, so a
// call never waits for the device work of the call before it.
struct Staging {
    struct Slot {
        void* host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    Slot slot[2];
    int next = 0;

    // a pinned buffer of >= bytes that no pending copy reads; *idx names it for copy()
    void* acquire(size_t bytes, int* idx)
    {
        Slot& s = slot[next];
        *idx = next;
        next ^= 1;
        if (s.pending) {
            HIP_CHECK(hipEventSynchronize(s.done));
            s.pending = false;
        }
        if (!s.done) HIP_CHECK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        if (s.cap < bytes) {
            if (s.host) HIP_CHECK(hipHostFree(s.host));
            s.host = nullptr;
            s.cap = 0;
            const size_t want = std::max<size_t>(bytes, 4096);
            if (hipHostMalloc(&s.host, want, 0) != hipSuccess) {
                (void)hipGetLastError();
                s.host = nullptr;
                throw ApiError(RFA_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
            }
            s.cap = want;
        }
        return s.host;
    }
    void copy(int idx, void* dst, size_t bytes, hipStream_t st)
    {
        HIP_CHECK(hipMemcpyAsync(dst, slot[idx].host, bytes, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipEventRecord(slot[idx].done, st));
        slot[idx].pending = true;
    }
    // after the streams are idle
    void release()
    {
        for (Slot& s : slot) {
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.host) (void)hipHostFree(s.host);
            s = Slot{};
        }
    }
};

// The pairs schedule's launch sequence as a captured HIP graph.  A call whose
// key (every host value the sequence bakes in: string and buffer addresses and
// sizes, kind, cutoffs, output) equals the previous eager call's is captured
// once; later calls with that key replay the graph.
struct PairsGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key, last_eager;
    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
};

inline uint64_t host_ceil_div(uint64_t a, uint64_t b) { return a / b + (a % b != 0); }

inline uint32_t log2_ceil(uint64_t x)
{
    uint32_t b = 0;
    while ((1ull << b) < x) ++b;
    return b;
}

template <typename F>
rfa_status guarded(F&& f)
{
    try {
        f();
        return RFA_OK;
    }
    catch (const ApiError& e) {
        return fail(e.st, e.what());
    }
    catch (const HipError& e) {
        return fail(RFA_ERR_HIP, e.what());
    }
    catch (const std::bad_alloc&) {
        return fail(RFA_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    catch (const std::exception& e) {
        return fail(RFA_ERR_INVALID_ARGUMENT, e.what());
    }
}

// Length-sorted, lane-interleaved copy of a string set for top-k scoring
// (1-byte units).  Built once per set on first use.
struct ServeIndex {
    uint32_t ngroups = 0;
    std::vector<uint32_t> glen_min, glen_max; // per group, over real slots
    uint32_t* d_data = nullptr;
    uint64_t* d_base = nullptr;
    uint32_t* d_dwords = nullptr;
    uint32_t* d_glen_min = nullptr;
    uint32_t* d_slot_len = nullptr;
    uint32_t* d_slot_idx = nullptr;
    uint64_t data_bytes = 0;
    ~ServeIndex();
};

} // namespace rfa

struct rfa_context {
    int device = 0;
    hipStream_t stream = nullptr;  // all results are ordered on this stream
    hipStream_t stream2 = nullptr; // second lane for work that overlaps (pairs buckets)
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex mu;
    int num_cus = 256;
    size_t lds_per_cu = 160 * 1024;
    rfa::DevBuf queue, pm_scratch, tflag, items_a, items_b, items_idx, items_swap, out_tmp, aux_tmp, misc;
    rfa::DevBuf lane_arena, lane_items, lane_swap, lane_metric, lane_out, block_items[4], plan_keys, plan_vals,
        plan_temp;
    rfa::DevBuf topk_scores, topk_run[2], topk_cand_s, topk_cand_i, topk_counts, topk_q, topk_out, qcut;
    rfa::DevBuf ed_items, ed_rows, ed_ops, band_q, band_w;
    rfa::DevBuf wl_bnd; // weighted Levenshtein strips: a boundary row per wave
    rfa::DevBuf ed_long; // Levenshtein editops past 64 units: a pair's rows per wave
    rfa::DevBuf ed_hnodes, ed_htmp; // Hirschberg splits: the pieces of one level, two rows per wave
    rfa::DevBuf pb_ws;   // partial ratios of needles over 1024 units: a workspace per wave
    rfa::DevBuf tset_cols, tset_rows; // partial_token_ratio: the repeat columns and rows
    size_t lane_arena_bytes_zeroed = 0;
    uint32_t gen_next = 1;
    rfa::Keep keep; // host sources of in-flight uploads
    rfa::Staging staging; // pinned per-call uploads of the pairs path
    rfa::PairsGraph pairs_graph;
    bool pairs_graph_on = false; // RFA_OPTION_PAIRS_GRAPH
    uint32_t seg_chain_steps = rfa::SEG_CHAIN_STEPS; // RFA_OPTION_PAIRS_CHAIN, set by rfa_pairs_autotune
    bool no_band = false; // RFA_OPTION_CDIST_BAND = 0
    // diagonal band kernel launch settings, set by rfa_cdist_autotune: words per lane for each
    // window-size bucket (band_bucket; 0: the rule of band_rule_wpl) and the workgroup size of
    // the launch (0: BAND_THREADS)
    uint8_t band_wpl[6] = {0, 0, 0, 0, 0, 0};
    uint16_t band_threads = 0;
    const double* pcut = nullptr; // per-pair cutoffs of the cdist launches in progress (WRatio)
    rfa::DevBuf wr_buf[3];         // WRatio's intermediate matrices
    uint64_t* d_counters = nullptr; // RFA_OPTION_COUNT_WORD_STEPS: band word-steps, band pairs
    void* noalloc_scratch = nullptr; // this device's static workspace (rfa_pairs_noalloc)
};

struct rfa_strings {
    rfa_context* ctx = nullptr;
    int width = 1;
    uint64_t count = 0;
    void* d_chars = nullptr;
    uint64_t* d_off = nullptr;
    std::vector<uint8_t> h_chars;
    std::vector<uint64_t> h_off;
    uint64_t max_len = 0;
    uint64_t id = 0;   // unique per set of the process (caches of derived sets key on it)
    bool view = false; // borrows d_chars/d_off of another set
    std::shared_ptr<rfa::ServeIndex> index; // built on first top-k use

    mutable int64_t max_unit_cached = -1;

    uint64_t len(uint64_t i) const { return h_off[i + 1] - h_off[i]; }
    // largest code unit of the set (views, which hold no host copy, report the width's maximum)
    uint32_t max_unit() const
    {
        if (max_unit_cached >= 0) return (uint32_t)max_unit_cached;
        uint32_t mx = 0;
        const uint64_t nunits = h_off.empty() ? 0 : h_off.back() - h_off.front();
        if (h_chars.size() < nunits * width)
            mx = width == 1 ? 0xffu : (width == 2 ? 0xffffu : 0xffffffffu);
        else if (width == 1)
            for (uint8_t c : h_chars) mx = c > mx ? c : mx;
        else if (width == 2) {
            const uint16_t* u = (const uint16_t*)h_chars.data();
            for (size_t i = 0, e = h_chars.size() / 2; i < e; ++i) mx = u[i] > mx ? u[i] : mx;
        }
        else {
            const uint32_t* u = (const uint32_t*)h_chars.data();
            for (size_t i = 0, e = h_chars.size() / 4; i < e; ++i) mx = u[i] > mx ? u[i] : mx;
        }
        max_unit_cached = mx;
        return mx;
    }
    uint32_t unit(uint64_t pos) const
    {
        if (width == 1) return h_chars[pos];
        if (width == 2) return ((const uint16_t*)h_chars.data())[pos];
        return ((const uint32_t*)h_chars.data())[pos];
    }
};

namespace rfa {

// deduplicated token-sorted copy of a string set and its token table (device)
struct TokenTable {
    rfa_strings* joined = nullptr; // the joins (owned unless handed over)
    uint8_t* d_blob = nullptr;
    uint64_t* d_off = nullptr;     // count + 1: the tokens of string i are [d_off[i], d_off[i + 1])
    uint64_t* d_hash = nullptr;
    uint32_t* d_start = nullptr;   // start in the string's join
    uint32_t* d_len = nullptr;
    ~TokenTable();
};

struct TileClass {
    uint32_t first_tile = 0, ntiles = 0;
    int word_bits = 64;
    uint32_t rows_cap = 2;
};

} // namespace rfa

struct rfa_scorer {
    rfa_context* ctx = nullptr;
    int metric = 0;
    rfa_weights w{1, 1, 1};
    double prefix_weight = 0.1;
    rfa_strings* q = nullptr; // owned copy of the queries
    // tile plan
    std::vector<uint32_t> tile_order; // query indices, 64 per tile
    uint32_t* d_tile_order = nullptr;
    std::vector<rfa::TileClass> classes;
    uint32_t hash_bits = 6;
    std::vector<uint32_t> block_queries; // scored by the block kernel
    uint32_t* d_block_queries = nullptr;
    uint64_t block_max_len = 0, block_min_len = 0;
    std::shared_ptr<void> few_plan; // cached plan of the one-pair-per-lane path (last choice set)
    uint32_t block_u8_rows = 0; // distinct symbols + 1 of the 8-bit block queries (0: unknown)
    std::vector<uint16_t> q_rows; // 8-bit queries: distinct symbols + 1 of each block query (others 0)
    // metric the caller asked for (metric holds the kernels' metric: token-sort
    // scorers run ratio / partial_ratio on token-sorted copies, QRatio runs ratio)
    int api_metric = 0;
    bool token_sort = false;
    bool empty_zero = false;
    bool signed8 = false;            // RFA_SCORER_SIGNED_UNITS
    rfa_strings* q_raw = nullptr;    // token-sort / token-set scorers: the queries as given
    // token-set scorers: the queries' token table (q holds their deduplicated joins), the
    // choices' (cached per set), and token_ratio's token-sort scorer
    bool token_set = false;
    std::shared_ptr<rfa::TokenTable> q_tokens;
    std::shared_ptr<rfa::TokenTable> c_tokens;
    uint64_t c_tokens_id = 0;
    bool c_tokens_signed8 = false;
    rfa_scorer* sort_sub = nullptr; // token_ratio: token-sort ratio; partial_token_ratio: partial token-sort
    rfa_scorer* pr_dedup = nullptr; // partial token ratios: partial_ratio over the deduplicated joins
    // partial_token_ratio: the queries that repeat a token (their deduplicated join is shorter
    // than their sorted join) and a partial_ratio scorer over just their deduplicated joins
    std::vector<uint32_t> dup_queries;
    uint32_t* d_dup_queries = nullptr; // dup_queries on the device (row scatter)
    rfa_scorer* pr_dedup_dupq = nullptr;
    bool dup_queries_ready = false;
    rfa_strings* tok_choices = nullptr; // token-sorted copy of the last choice set
    // fuzz::WRatio: its sub-scorers over the raw queries (ratio, token_ratio, partial_ratio,
    // partial_token_ratio), run as whole cdist launches with per-pair raised cutoffs
    bool wratio = false;
    rfa_scorer* wr_sub[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t tok_choices_id = 0;
    bool tok_choices_signed8 = false;
};

namespace rfa {

void begin_call(rfa_context* ctx);
int algo_for(int metric, const rfa_weights& w);
bool multi_levenshtein_weights(const rfa_weights& w);
bool is_int_kind(int metric, int kind);
rfa_strings* make_strings(rfa_context* ctx, const void* chars, int width, const uint64_t* off, uint64_t count);
// token-sorted copy of a set (device transform on the context's stream)
rfa_strings* token_sorted(rfa_context* ctx, const rfa_strings* in, bool signed8);
// deduplicated token-sorted copy and token table
std::shared_ptr<rfa::TokenTable> token_table(rfa_context* ctx, const rfa_strings* in, bool signed8);
// the choice set a scorer's kernels read: its token-sorted copy for token-sort scorers
const rfa_strings* scorer_choices(rfa_scorer* sc, const rfa_strings* ch);
void free_strings(rfa_strings* s);
rfa_scorer* make_scorer(rfa_context* ctx, int metric, const rfa_strings* queries, const rfa_weights& w,
                        double prefix_weight);
void free_scorer(rfa_scorer* sc);
void validate_kind(int metric, int kind, int semantics);
// qcut_host: optional per-query integer cutoffs (host memory, one per query)
void cdist_tset(rfa_scorer* sc, const rfa_strings* raw, int kind, int semantics, double cd, void* out,
                int out_on_device);
void cdist_wratio(rfa_scorer* sc, const rfa_strings* raw, int kind, int semantics, double cd, void* out,
                  int out_on_device);
void cdist_impl(rfa_scorer* sc, const rfa_strings* ch, int kind, int semantics, uint64_t cu, double cd, void* out,
                int out_on_device, const uint64_t* qcut_host = nullptr);

// rfa_topk.cpp
// cdist of at most 16 one-byte queries (<= 256 units, LCS or uniform Levenshtein
// recurrence) against a set, one pair per lane; false when it does not apply.
bool cdist_few_queries(rfa_scorer* sc, const rfa_strings* ch, const ScoreSpec& spec, int algo, void* dout,
                       const uint64_t* d_qcut = nullptr);
void topk_impl(rfa_scorer* sc, const rfa_strings* ch, int kind, double cutoff, uint32_t k, int64_t* out_index,
               double* out_score);

} // namespace rfa
