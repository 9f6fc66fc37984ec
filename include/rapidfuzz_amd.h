/*
 * rapidfuzz_amd.h -- C ABI of the MI355X (gfx950) bit-parallel distance engine.
 *
 * The reference (rapidfuzz-cpp 3.3.3) is a header-only C++ template library
 * with no C FFI; its "operator API" is the scorer classes.  Each entry point
 * below names the reference interface whose work it takes over.  Plain
 * pointers and sizes only; no torch or HIP types cross this boundary (the
 * stream is handed out as an opaque pointer).
 *
 * Strings are sequences of unsigned code units of width 1, 2 or 4 bytes,
 * passed packed: chars[offsets[i] .. offsets[i+1]) is string i.
 *
 * Threading: a context owns one HIP stream; calls on one context are
 * serialized by an internal mutex, calls on different contexts may run
 * concurrently.  Scorers and string sets are immutable after creation, like
 * the reference's const scorer objects (README.md:216-236).
 *
 * Errors: every call returns an rfa_status; rfa_last_error() gives the
 * message of the last failed call on the calling thread.  Messages reuse the
 * reference's exception texts where the reference throws
 * (details/distance.hpp:426-427, distance/Levenshtein.hpp:338-339,372).
 */
#ifndef RAPIDFUZZ_AMD_H
#define RAPIDFUZZ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFA_ABI_VERSION 1

typedef enum {
    RFA_OK = 0,
    RFA_ERR_INVALID_ARGUMENT = 1, /* the reference throws std::invalid_argument */
    RFA_ERR_UNSUPPORTED = 2,      /* outside the accelerated path (e.g. size limits of a kernel) */
    RFA_ERR_HIP = 3,              /* HIP runtime error */
    RFA_ERR_OUT_OF_MEMORY = 4,
    RFA_ERR_NO_DEVICE = 5
} rfa_status;

/* Metric of a scorer (the reference's scorer classes). */
typedef enum {
    RFA_LEVENSHTEIN = 0,  /* CachedLevenshtein / experimental::MultiLevenshtein */
    RFA_INDEL = 1,        /* CachedIndel / experimental::MultiIndel */
    RFA_LCSSEQ = 2,       /* CachedLCSseq / experimental::MultiLCSseq */
    RFA_OSA = 3,          /* CachedOSA / experimental::MultiOSA */
    RFA_JARO = 4,         /* CachedJaro / experimental::MultiJaro */
    RFA_JARO_WINKLER = 5, /* CachedJaroWinkler / experimental::MultiJaroWinkler */
    RFA_RATIO = 6,        /* fuzz::CachedRatio / fuzz::experimental::MultiRatio */
    RFA_HAMMING = 7,      /* CachedHamming (Hamming.hpp:131-161) with pad = true; Cached semantics only */
    RFA_PREFIX = 8,       /* CachedPrefix (Prefix.hpp:65-94); Cached semantics only */
    RFA_POSTFIX = 9,      /* CachedPostfix (Postfix.hpp:65-94); Cached semantics only */
    RFA_PARTIAL_RATIO = 10, /* fuzz::CachedPartialRatio (fuzz.hpp:169-200, fuzz_impl.hpp:57-279): best ratio of
                               the shorter string against the longer one's windows; similarity, Cached only */
    /* ratio / partial_ratio of the token-sorted strings (each string split at whitespace, tokens
     * sorted, joined with single spaces: detail::sorted_split(...).join()); similarity, Cached only */
    RFA_TOKEN_SORT_RATIO = 11,         /* fuzz::CachedTokenSortRatio (fuzz.hpp:282-304) */
    RFA_PARTIAL_TOKEN_SORT_RATIO = 12, /* fuzz::CachedPartialTokenSortRatio (fuzz.hpp:342-364) */
    RFA_QRATIO = 13,                   /* fuzz::CachedQRatio (fuzz.hpp:776-796): ratio, 0 if either string is empty */
    /* token sets (deduplicated sorted tokens): rfa_cdist only; queries whose deduplicated token join
     * has at most 4096 code units */
    RFA_TOKEN_SET_RATIO = 14,          /* fuzz::CachedTokenSetRatio (fuzz.hpp:412-434) */
    RFA_TOKEN_RATIO = 15,              /* fuzz::CachedTokenRatio (fuzz.hpp:530-557): max of the token-sort and token-set ratios */
    RFA_PARTIAL_TOKEN_SET_RATIO = 16,  /* fuzz::CachedPartialTokenSetRatio (fuzz.hpp:471-493) */
    RFA_PARTIAL_TOKEN_RATIO = 17,      /* fuzz::CachedPartialTokenRatio (fuzz.hpp:595-620) */
    RFA_DAMERAU_LEVENSHTEIN = 18,      /* experimental::CachedDamerauLevenshtein (DamerauLevenshtein.hpp:112-140):
                                          Cached semantics, every kind; strings below 2^29 code units */
    RFA_WRATIO = 19                    /* fuzz::CachedWRatio (fuzz.hpp:659-682, fuzz_impl.hpp:790-876): ratio, then
                                          token_ratio or partial_ratio and partial_token_ratio with cutoffs raised per
                                          pair; similarity, Cached only; the token ratios' and partial_ratio's
                                          limits apply */
} rfa_metric;

/* Scorer method (details/distance.hpp:211-301, 326-356). */
typedef enum {
    RFA_DISTANCE = 0,
    RFA_SIMILARITY = 1,
    RFA_NORMALIZED_DISTANCE = 2,
    RFA_NORMALIZED_SIMILARITY = 3
} rfa_kind;

/* Which reference class's result semantics to reproduce. */
typedef enum {
    RFA_CACHED = 0, /* Cached*: cutoff applied inside (details/distance.hpp:210-388) */
    RFA_MULTI = 1   /* experimental::Multi*: full score, cutoff after (details/distance.hpp:390-582) */
} rfa_semantics;

/* Pair schedules for rfa_pairs (the uneven batch). */
typedef enum {
    RFA_SCHEDULE_BUCKETED = 0, /* length-bucketed persistent launch (the default) */
    RFA_SCHEDULE_PLAIN = 1     /* one pair per lane, all words sequential in the lane */
} rfa_schedule;

/* LevenshteinWeightTable (details/types.hpp:19-23). */
typedef struct {
    uint64_t insert_cost;
    uint64_t delete_cost;
    uint64_t replace_cost;
} rfa_weights;

typedef struct rfa_context rfa_context;
typedef struct rfa_strings rfa_strings;
typedef struct rfa_scorer rfa_scorer;

/* ---- context ------------------------------------------------------------ */

/* One HIP stream plus memory on `device`.  No reference counterpart: the
 * reference runs inline on the caller's thread. */
rfa_status rfa_context_create(int device, rfa_context** out);
rfa_status rfa_context_destroy(rfa_context* ctx);
rfa_status rfa_context_synchronize(rfa_context* ctx);
/* The context's hipStream_t, for callers that time or order work. */
void* rfa_context_stream(rfa_context* ctx);
/* Context options (no reference counterpart).
 * RFA_OPTION_PAIRS_GRAPH (default 0): 1 = an rfa_pairs call (bucketed schedule)
 * identical to the previous one (same string sets, kind, cutoffs, output and
 * buffer sizes) replays its launch sequence from a captured HIP graph. */
#define RFA_OPTION_PAIRS_GRAPH 1
/* RFA_OPTION_PAIRS_CHAIN (default 512, in [1, 2^31]): in the bucketed pairs
 * schedule, a pair whose longer string has this many units or more (up to
 * 4096) uses the longer string as the bit-parallel pattern (lanes per pair =
 * its words), and pairs whose text has this many units or more are scheduled
 * first.  rfa_pairs_autotune sets it by measurement. */
#define RFA_OPTION_PAIRS_CHAIN 2
/* RFA_OPTION_CDIST_BAND (default 1): cdist of 8-bit queries over 64 units under a
 * Levenshtein distance cutoff k of at most 4094 runs the diagonal band kernel
 * (a window of ceil((k + 2) / 64) words per pair, several pairs per wave); 0 =
 * the one-wave-per-pair block kernel for every long query (same results). */
#define RFA_OPTION_CDIST_BAND 3
/* RFA_OPTION_COUNT_WORD_STEPS (default 0): 1 = the band kernel counts, on the
 * device, the 64-bit word-steps it evaluates (text units run x window words of
 * every pair) and the pairs it runs, for rfa_context_counter; setting it
 * (to 0 or 1) clears both counters. */
#define RFA_OPTION_COUNT_WORD_STEPS 4
rfa_status rfa_context_set_option(rfa_context* ctx, int option, int64_t value);
/* Device counters of the context (no reference counterpart): waits for the
 * context's work, then stores the counter in *value. */
#define RFA_COUNTER_BAND_WORD_STEPS 1
#define RFA_COUNTER_BAND_PAIRS 2
rfa_status rfa_context_counter(rfa_context* ctx, int counter, uint64_t* value);
/* Message of the last failed call on this thread ("" if none). */
const char* rfa_last_error(void);
int rfa_abi_version(void);

/* ---- device-resident string sets ---------------------------------------- */

/* Copies `count` packed strings to the device.  char_width is 1, 2 or 4.
 * Replaces the per-call Range<InputIt> views of s2 (details/Range.hpp:73-186)
 * with one resident set scored many times. */
rfa_status rfa_strings_create(rfa_context* ctx, const void* chars, int char_width, const uint64_t* offsets,
                              uint64_t count, rfa_strings** out);
rfa_status rfa_strings_destroy(rfa_strings* set);
uint64_t rfa_strings_count(const rfa_strings* set);

/* ---- scorers ------------------------------------------------------------ */

/* A scorer holding one cached query per string of `queries` -- the shared
 * state of Cached<Metric><CharT>(s1) (e.g. distance/Levenshtein.hpp:424-432,
 * Indel.hpp:149-159, LCSseq.hpp:197-206, OSA.hpp:232-241,
 * JaroWinkler.hpp:168-178, fuzz.hpp:99-118) and of
 * experimental::Multi<Metric><MaxLen>(count) after count inserts
 * (Levenshtein.hpp:334-380).  weights may be NULL ({1,1,1}); only Levenshtein
 * reads it.  prefix_weight is read by Jaro-Winkler only.
 * Levenshtein takes any weights (CachedLevenshtein::_distance,
 * Levenshtein.hpp:445-477): weights other than uniform ({w,w,w}) or Indel-like
 * (ins == del, replace >= ins + del) run the generalized Wagner-Fischer
 * recurrence (Levenshtein_impl.hpp:85-155), where rfa_cdist returns
 * RFA_ERR_UNSUPPORTED for a weight of 2^24 or more, for (query length +
 * choice length + 1) x largest weight of 2^30 or more, and for pairs in
 * which both strings exceed 1,024 code units. */
rfa_status rfa_scorer_create(rfa_context* ctx, int metric, const rfa_strings* queries, const rfa_weights* weights,
                             double prefix_weight, rfa_scorer** out);
rfa_status rfa_scorer_destroy(rfa_scorer* scorer);
/* Scorer options (no reference counterpart).
 * RFA_SCORER_SIGNED_UNITS (default 0): 1 = 8-bit code units order as signed values when
 * tokens are sorted (the reference's comparison for `char` strings, Range.hpp:214-218);
 * 0 = as unsigned values (uint8_t strings, and every 16/32-bit width). */
#define RFA_SCORER_SIGNED_UNITS 1
rfa_status rfa_scorer_set_option(rfa_scorer* scorer, int option, int64_t value);

/* All queries against all choices: out[q * n_choices + c] is
 *   Cached<Metric>(query q).<kind>(choice c, cutoff)        (semantics RFA_CACHED)
 *   Multi<Metric>.<kind>(scores, ..., choice c, cutoff)[q]  (semantics RFA_MULTI)
 * Integer kinds of Levenshtein/Indel/LCSseq/OSA write uint64_t with
 * cutoff_u (INT64_MAX = the reference default); every other combination
 * writes double with cutoff_d.  `out` is a device pointer when
 * out_on_device != 0, else host memory.  Replaces the caller's loop of
 * scorer.distance(s2) calls (README.md:153-214; distance/Levenshtein.hpp:383-399
 * for the Multi* per-text call). */
rfa_status rfa_cdist(rfa_scorer* scorer, const rfa_strings* choices, int kind, int semantics, uint64_t cutoff_u,
                     double cutoff_d, void* out, int out_on_device);

/* As rfa_cdist for the integer kinds, with one cutoff per query:
 *   out[q * n_choices + c] = Cached<Metric>(query q).<kind>(choice c, cutoffs_u[q])
 * (or the Multi* value cut at cutoffs_u[q]).  This is the reference's usage loop
 * when every query has its own cutoff (e.g. a maximum distance proportional to
 * its length), done as one call so the pairs of all queries run together.
 * `cutoffs_u` is host memory with one entry per query of the scorer. */
rfa_status rfa_cdist_cutoffs(rfa_scorer* scorer, const rfa_strings* choices, int kind, int semantics,
                             const uint64_t* cutoffs_u, void* out, int out_on_device);

/* Independent pairs: out[i] = metrics[i](a[i], b[i]).<kind>(cutoff) with
 * Cached* semantics; metrics[i] is one of Levenshtein (weights {1,1,1}),
 * Indel, LCSseq, OSA.  Replaces per-pair calls of the free functions
 * levenshtein_distance / indel_distance (Levenshtein.hpp:140-265,
 * Indel.hpp:20-60).  `metrics` is host memory; a and b must hold the same
 * number of strings and the same char width. */
rfa_status rfa_pairs(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const int32_t* metrics, int kind,
                     uint64_t cutoff_u, double cutoff_d, int schedule, void* out, int out_on_device);
/* rfa_pairs without any allocation (no reference counterpart): for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  `metrics` (one code per
 * pair: RFA_LEVENSHTEIN, RFA_INDEL, RFA_LCSSEQ or RFA_OSA) and `out` (8 bytes
 * per pair) are device memory; the call only enqueues kernels on the context's
 * stream (rfa_context_stream): no allocation, no copy, no synchronization.
 * The bucketed schedule's plan lives in a static per-GPU workspace of 48 MiB
 * held from library load to process end, so one call takes at most 2^20 pairs,
 * and every pair's shorter string must have at most 4096 units (64 when
 * has_osa is nonzero: `metrics` may then hold RFA_OSA); larger calls fail with
 * RFA_ERR_UNSUPPORTED.  Codes outside 0..3 give unspecified values.  Calls that
 * overlap on one GPU (from other contexts or threads without ordering) share
 * that workspace and race; calls on one context are ordered by its stream. */
rfa_status rfa_pairs_noalloc(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const uint8_t* metrics,
                             int has_osa, int kind, uint64_t cutoff_u, double cutoff_d, void* out);
/* Autotuner of the bucketed pairs schedule (no reference counterpart): times
 * the schedule on this batch for each candidate RFA_OPTION_PAIRS_CHAIN value
 * (128 .. 2048, and 4097 = the shorter string is always the pattern), keeps the
 * fastest in the context and returns it in *chosen.  Results do not depend on
 * the choice; it runs the batch 24 times. */
rfa_status rfa_pairs_autotune(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, const int32_t* metrics,
                              int kind, uint64_t cutoff_u, double cutoff_d, uint32_t* chosen);

/* Autotuner of the diagonal band kernel (no reference counterpart: the reference's
 * banded recurrence, Levenshtein_impl.hpp:641-850, has no launch settings).  Runs the
 * cdist of the scorer's queries over 64 units against `choices` under their Levenshtein
 * distance cutoffs (cutoffs_u per query, or cutoff_u for all when NULL) and measures, for
 * each bucket of window sizes (band words 1, 2-3, 4-7, 8-15, 16-31, 32-64), the words per
 * lane (1 or 2: lanes per pair = ceil(band words / words per lane), so the pairs one wave
 * runs), then for the whole launch the workgroup size (64, 128 or 256 threads: the pairs a
 * workgroup runs against its query); keeps the fastest in the context for later cdist
 * calls and writes them to chosen[2 * bucket] (words per lane) and chosen[2 * bucket + 1]
 * (threads), 0 for buckets without such queries (chosen may be NULL).  Results do not
 * depend on the settings. */
rfa_status rfa_cdist_autotune(rfa_scorer* sc, const rfa_strings* choices, int kind, uint64_t cutoff_u,
                              const uint64_t* cutoffs_u, uint32_t* chosen);

/* Top-k extraction.  Replaces the README's extract() loop (README.md:153-214):
 * for every query, score = Cached<Metric>(query).<kind>(choice, cutoff_d) for
 * every choice, results with score >= cutoff_d, ordered by (score descending,
 * choice index ascending), first k kept.  `kind` must be a double similarity:
 * RFA_NORMALIZED_SIMILARITY, or RFA_SIMILARITY of Jaro, Jaro-Winkler or
 * ratio.  Entries beyond the number of qualifying choices get index -1 and
 * score 0.  out_index[q * k + j], out_score[q * k + j]: host memory. */
rfa_status rfa_topk(rfa_scorer* scorer, const rfa_strings* choices, int kind, double cutoff_d, uint32_t k,
                    int64_t* out_index, double* out_score);

/* ---- edit operations ---------------------------------------------------------
 * rfa_editops: the edit script of every pair (a[i], b[i]), as
 * levenshtein_editops(a[i], b[i]) (metric RFA_LEVENSHTEIN; Levenshtein.hpp:282-295,
 * Levenshtein_impl.hpp:1229-1257) or lcs_seq_editops / indel_editops (RFA_LCSSEQ or
 * RFA_INDEL; LCSseq.hpp:65-75, Indel.hpp:65-75) or hamming_editops with pad (RFA_HAMMING;
 * Hamming.hpp:77-90).  Operations have the reference's
 * EditOp layout (type, src_pos, dest_pos; details/types.hpp:28-60); pair i's run
 * is ops[offsets[i], offsets[i + 1]), its src_len / dest_len are the string lengths.
 * offsets (count + 1 entries) is always written.  ops may be NULL to ask for the
 * sizes only; otherwise capacity (in operations) must be at least offsets[count]
 * (else RFA_ERR_INVALID_ARGUMENT, with offsets written).  A pair whose first string
 * has more than 4096 code units once the common prefix and suffix are removed
 * returns RFA_ERR_UNSUPPORTED.  Levenshtein pairs the reference splits by
 * Hirschberg (stripped |s1| > 64, |s2| >= 10 and |s1| * |s2| >= 2^22,
 * Levenshtein_impl.hpp:1185-1212) are split the same way; they need string sets
 * that hold a host copy (views return RFA_ERR_UNSUPPORTED for them).  Hamming has
 * no limit. */
#define RFA_EDIT_NONE 0
#define RFA_EDIT_REPLACE 1
#define RFA_EDIT_INSERT 2
#define RFA_EDIT_DELETE 3
typedef struct {
    int32_t type; /* RFA_EDIT_* */
    int32_t reserved;
    uint64_t src_pos;
    uint64_t dest_pos;
} rfa_editop;
rfa_status rfa_editops(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b, int metric, uint64_t* offsets,
                       rfa_editop* ops, uint64_t capacity);

/* ---- partial ratio alignment -------------------------------------------------
 * rfa_partial_ratio_alignment: fuzz::partial_ratio_alignment(a[i], b[i], score_cutoff)
 * (fuzz.hpp:129-135, fuzz_impl.hpp:196-230) of every pair: the partial ratio and the
 * windows of a[i] (src) and b[i] (dest) it was found at, with the reference's choice
 * among equal windows.  The layout is ScoreAlignment<double>'s (details/types.hpp:567-594).
 * out: count entries in host memory.  A pair whose shorter string has more than 4096
 * code units returns RFA_ERR_UNSUPPORTED. */
typedef struct {
    double score;
    uint64_t src_start;
    uint64_t src_end;
    uint64_t dest_start;
    uint64_t dest_end;
} rfa_score_alignment;
rfa_status rfa_partial_ratio_alignment(rfa_context* ctx, const rfa_strings* a, const rfa_strings* b,
                                       double score_cutoff, rfa_score_alignment* out);

/* ---- request server (continuous batching) -------------------------------
 *
 * No reference counterpart: the reference leaves the extract() loop to the
 * caller (README.md:153-214).  A server holds a device-resident copy of a
 * choice set and answers top-k requests (as rfa_topk, one query each).
 * Requests queue on submit; each of `slots` worker threads, whenever its
 * previous batch is done, admits the pending requests in arrival order (up
 * to max_batch; requests for the fused kernel and for the dense path go in
 * separate batches) and runs them on its own HIP stream, so batches of
 * different slots overlap on the device.  A request's result does not depend
 * on the batch it ran in. */
typedef struct rfa_server rfa_server;

typedef struct {
    int metric;             /* rfa_metric */
    int kind;               /* a double similarity kind, as for rfa_topk */
    double score_cutoff;
    uint32_t k;
    uint32_t max_batch;     /* requests admitted into one batch, 1..4096 */
    uint32_t slots;         /* batches in flight at once, 1..8 */
    uint32_t candidate_cap; /* results >= cutoff kept per request before ordering; a request
                               with more is recomputed through the dense path */
    double prefix_weight;   /* Jaro-Winkler */
    rfa_weights weights;    /* Levenshtein; all zero means {1,1,1} */
} rfa_server_config;

typedef struct {
    uint64_t requests;
    uint64_t batches;
    uint64_t max_batch_seen;
    uint64_t dense_fallbacks;
} rfa_server_stats;

rfa_status rfa_server_create(rfa_context* ctx, const rfa_strings* choices, const rfa_server_config* cfg,
                             rfa_server** out);
/* Finishes the pending requests, then frees the server. */
rfa_status rfa_server_destroy(rfa_server* srv);
/* Queues one query (len code units of the choice set's width); returns at once. */
rfa_status rfa_server_submit(rfa_server* srv, const void* query, uint64_t len, uint64_t* ticket);
/* Blocks until the request is answered; k entries each (either may be NULL). */
rfa_status rfa_server_wait(rfa_server* srv, uint64_t ticket, int64_t* out_index, double* out_score);
/* Load replay: submits query i at arrival_s[i] seconds after the call starts
 * (from a feeder thread), waits for all, and writes latency_s[i] = answer
 * time - scheduled arrival.  Results as rfa_server_wait (k per query; the
 * output pointers may be NULL). */
rfa_status rfa_server_replay(rfa_server* srv, const void* chars, const uint64_t* offsets, uint64_t count,
                             const double* arrival_s, double* latency_s, int64_t* out_index, double* out_score);
rfa_status rfa_server_get_stats(rfa_server* srv, rfa_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* RAPIDFUZZ_AMD_H */
