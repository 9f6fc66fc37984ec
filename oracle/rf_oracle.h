/*
 * rf_oracle.h -- CPU restatement of rapidfuzz-cpp's scoring semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code: it exists so that tests/, the smoke check and the CPU-baseline
 * leg of bench.py have an independent checker for the HIP kernels.
 *
 * Two layers:
 *   1. raw values by naive O(m*n) dynamic programming, restated from the
 *      reference's own naive oracle (the rapidfuzz_reference directory);
 *   2. the API layer (distance / similarity / normalized_* with score_cutoff)
 *      restated operation by operation from rapidfuzz/details/distance.hpp,
 *      rapidfuzz/details/common.hpp:50-53 and the per-metric Cached* classes.
 *
 * All strings are passed as uint32_t code units.  Build with
 * -ffp-contract=off so that double formulas round exactly like the
 * reference built with the same flag.
 */
#ifndef RF_ORACLE_H
#define RF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_LEVENSHTEIN = 0,
    ORC_INDEL = 1,
    ORC_LCSSEQ = 2,
    ORC_OSA = 3,
    ORC_JARO = 4,
    ORC_JARO_WINKLER = 5,
    ORC_RATIO = 6,
    ORC_HAMMING = 7, /* CachedHamming with pad = true */
    ORC_PREFIX = 8,
    ORC_POSTFIX = 9,
    ORC_PARTIAL_RATIO = 10, /* fuzz::CachedPartialRatio: similarity only */
    ORC_TOKEN_SORT_RATIO = 11,
    ORC_PARTIAL_TOKEN_SORT_RATIO = 12,
    ORC_QRATIO = 13,
    ORC_TOKEN_SET_RATIO = 14,
    ORC_TOKEN_RATIO = 15,
    ORC_PARTIAL_TOKEN_SET_RATIO = 16,
    ORC_PARTIAL_TOKEN_RATIO = 17,
    ORC_DAMERAU_LEVENSHTEIN = 18, /* experimental::CachedDamerauLevenshtein */
    ORC_WRATIO = 19               /* fuzz::CachedWRatio */
};

enum { ORC_DISTANCE = 0, ORC_SIMILARITY = 1, ORC_NORM_DISTANCE = 2, ORC_NORM_SIMILARITY = 3 };

/* ---- raw values (naive DP) ---- */
uint64_t orc_levenshtein_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2,
                             uint64_t ins, uint64_t del, uint64_t rep);
/* The LCS length by the bit-parallel recurrence (the Indel / LCS scorers' raw value) */
uint64_t orc_lcs_bits(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
uint64_t orc_lcs_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
uint64_t orc_osa_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
uint64_t orc_damerau_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
/* Editops of levenshtein_editops (metric 0) / lcs_seq_editops = indel_editops (metric 1
 * or 2) / hamming_editops with pad (metric 7) for one pair; writes at most l1 + l2
 * operations, returns their count. */
/* Levenshtein editops past 64 stripped units: 1 the reference's block band (default), 0 the
 * full-matrix walk */
void orc_set_lev_block(int on);
size_t orc_editops(int metric, const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, int32_t* type,
                   uint64_t* src_pos, uint64_t* dest_pos);
/* positional metrics: Hamming distance with padding, common prefix / suffix length */
uint64_t orc_hamming_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
uint64_t orc_prefix_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
uint64_t orc_postfix_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2);
void orc_jaro_raw(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl, uint64_t* common,
                  uint64_t* transpositions);
double orc_jaro_similarity(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl, double cutoff);
double orc_jaro_winkler_similarity(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl,
                                   double prefix_weight, double cutoff);

/* Code units of the token scorers: width 1, 2 or 4 (1 uses the reference's 8-bit whitespace
 * set) and, for width 1, signed comparison (the reference's `char`).  Default 4, unsigned. */
void orc_set_units(int width, int signed8);
/* detail::sorted_split(s).join() into out (>= n units); returns its length */
size_t orc_sorted_join(const uint32_t* s, size_t n, uint32_t* out);

/* fuzz::partial_ratio(s1, s2, cutoff): the reference's window search, step by step */
double orc_partial_ratio(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff);
/* fuzz::partial_ratio_alignment(s1, s2, cutoff): *score and al = {src_start, src_end, dest_start, dest_end} */
void orc_partial_ratio_alignment(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff,
                                 double* score, size_t* al);

/* ---- API layer: Cached<Metric>(s1).<kind>(s2, cutoff) ----
 * Returns 0 and writes *out_u for integer results, 1 and writes *out_d for
 * double results, -1 for an unsupported (metric, kind). */
int orc_cached_score(int metric, int kind, const uint32_t* s1, size_t l1, const uint32_t* s2,
                     size_t l2, uint64_t ins, uint64_t del, uint64_t rep, double prefix_weight,
                     uint64_t cutoff_u, double cutoff_d, uint64_t* out_u, double* out_d);

/* ---- API layer: experimental::Multi<Metric><MaxLen> for one inserted s1 ----
 * (the Multi* classes score unbounded first and apply the cutoff after).
 * MultiLevenshtein accepts weights {1, 1, <= 2} only (Levenshtein.hpp:334-340,
 * else -1); its distance is the uniform one whatever the weights
 * (Levenshtein.hpp:383-399) while maximum() uses them (:401-404). */
int orc_multi_score(int metric, int kind, const uint32_t* s1, size_t l1, const uint32_t* s2,
                    size_t l2, uint64_t ins, uint64_t del, uint64_t rep, double prefix_weight,
                    uint64_t cutoff_u, double cutoff_d, uint64_t* out_u, double* out_d);

/* ---- batch helpers (single thread), out[q * nc + c] ----
 * chars: packed code units; offsets: count+1 entries. */
int orc_cdist(int metric, int kind, int multi, const uint32_t* qchars, const uint64_t* qoff,
              size_t nq, const uint32_t* cchars, const uint64_t* coff, size_t nc, uint64_t ins,
              uint64_t del, uint64_t rep, double prefix_weight, uint64_t cutoff_u, double cutoff_d,
              uint64_t* out_u, double* out_d);

/* pairs: out[i] = score(a[i], b[i]) with metric[i] */
int orc_pairs(const int32_t* metrics, int kind, const uint32_t* achars, const uint64_t* aoff,
              const uint32_t* bchars, const uint64_t* boff, size_t n, uint64_t cutoff_u,
              double cutoff_d, uint64_t* out_u, double* out_d);

#ifdef __cplusplus
}
#endif
#endif
