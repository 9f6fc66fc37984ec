// rfa_semantics.hpp -- device-side result semantics (score_cutoff, normalization).
//
// Every kernel computes an exact raw value per pair (edit distance, LCS
// length, or the Jaro common/transposition counts).  The functions here turn
// it into the value the reference API returns, following the reference's
// CRTP bases operation by operation so integer results are identical and
// double results round identically:
//
//   Cached* distance/similarity ........ rapidfuzz/details/distance.hpp:269-388
//   Cached* normalized_* ................ rapidfuzz/details/distance.hpp:210-267
//   NormSim_to_NormDist (+1e-5) ......... rapidfuzz/details/common.hpp:50-53
//   experimental::Multi* (cut after) .... rapidfuzz/details/distance.hpp:390-582
//   Levenshtein weights dispatch ........ rapidfuzz/distance/Levenshtein.hpp:445-477
//   Jaro filters and formula ............ rapidfuzz/distance/Jaro_impl.hpp:32-71,409-454
//   Jaro-Winkler cutoff and boost ....... rapidfuzz/distance/JaroWinkler_impl.hpp:41-71
//   MultiJaroWinkler .................... rapidfuzz/distance/JaroWinkler.hpp:122-145
//   fuzz::ratio / CachedRatio ........... rapidfuzz/fuzz_impl.hpp:28-47, fuzz.hpp:82-89
//   fuzz::QRatio / CachedQRatio ......... rapidfuzz/fuzz_impl.hpp:897-922
//
// Compile with -ffp-contract=off: the JW boost must not become an FMA.
#pragma once
#include <stdint.h>

#include "rfa_types.hpp"

namespace rfa {

struct RawResult {
    uint64_t v;      // distance (Levenshtein/OSA), LCS length (Indel/LCSseq/ratio), common chars (Jaro)
    uint32_t tr;     // Jaro: mismatched flagged pairs (halved inside the formula)
    uint32_t prefix; // Jaro-Winkler: common prefix length, at most 4
};

__device__ __forceinline__ uint64_t u64min(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint64_t u64max(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t ceil_div_u64(uint64_t a, uint64_t b) { return a / b + (uint64_t)(a % b != 0); }

// --- integer metrics -------------------------------------------------------

struct IntPair {
    const ScoreSpec* s;
    uint64_t l1, l2;
    uint64_t raw; // exact distance (LEV uniform, OSA) or exact LCS (INDEL, LCSSEQ, LEV indel-weights)
};

__device__ __forceinline__ uint64_t int_maximum(const IntPair& p)
{
    const ScoreSpec& s = *p.s;
    switch (s.metric) {
    case RFA_LEVENSHTEIN: {
        // Levenshtein_impl.hpp:118-128
        uint64_t m = p.l1 * s.del + p.l2 * s.ins;
        if (p.l1 >= p.l2)
            m = u64min(m, p.l2 * s.rep + (p.l1 - p.l2) * s.del);
        else
            m = u64min(m, p.l1 * s.rep + (p.l2 - p.l1) * s.ins);
        return m;
    }
    case RFA_INDEL: return p.l1 + p.l2;
    default: return u64max(p.l1, p.l2); // LCSseq, OSA
    }
}

// similarity-type integer metrics (CachedSimilarityBase): the raw value is the similarity
__device__ __forceinline__ bool is_sim_int(int32_t m) { return m == RFA_LCSSEQ || m == RFA_PREFIX || m == RFA_POSTFIX; }

// exact Indel distance from the LCS length
__device__ __forceinline__ uint64_t indel_from_lcs(const IntPair& p) { return p.l1 + p.l2 - 2 * p.raw; }

// derived._distance(s2, c) of the distance metrics, from the exact value.
__device__ __forceinline__ uint64_t int_distance_core(const IntPair& p, uint64_t c)
{
    const ScoreSpec& s = *p.s;
    if (s.metric == RFA_LEVENSHTEIN) {
        // general weights: the exact weighted distance, cut (Levenshtein_impl.hpp:141-152)
        if (s.algo == ALGO_WLEV) return p.raw <= c ? p.raw : c + 1;
        // CachedLevenshtein::_distance: uniform or Indel, scaled by the common weight
        if (s.ins == s.del) {
            if (s.ins == 0) return 0;
            uint64_t nc = s.ins == 1 ? c : ceil_div_u64(c, s.ins);
            uint64_t d;
            if (s.ins == s.rep) {
                nc = u64min(nc, u64max(p.l1, p.l2)); // uniform_levenshtein_distance clamp
                d = p.raw <= nc ? p.raw : nc + 1;
            }
            else {
                uint64_t exact = indel_from_lcs(p);
                d = exact <= nc ? exact : nc + 1;
            }
            d *= s.ins;
            return d <= c ? d : c + 1;
        }
        return c + 1; // unreachable: other weights run ALGO_WLEV
    }
    uint64_t exact = (s.metric == RFA_INDEL) ? indel_from_lcs(p) : p.raw;
    return exact <= c ? exact : c + 1;
}

// LCSseq, Prefix and Postfix are similarity metrics: CachedSimilarityBase::_distance
__device__ __forceinline__ uint64_t lcs_distance(const IntPair& p, uint64_t c)
{
    uint64_t maximum = int_maximum(p);
    uint64_t cs = (maximum > c) ? maximum - c : 0;
    uint64_t sim = p.raw >= cs ? p.raw : 0;
    uint64_t dist = maximum - sim;
    return dist <= c ? dist : c + 1;
}

__device__ __forceinline__ uint64_t int_api_distance(const IntPair& p, uint64_t c)
{
    if (is_sim_int(p.s->metric)) return lcs_distance(p, c);
    return int_distance_core(p, c);
}

__device__ __forceinline__ uint64_t int_api_similarity(const IntPair& p, uint64_t c)
{
    if (is_sim_int(p.s->metric)) return p.raw >= c ? p.raw : 0;
    uint64_t maximum = int_maximum(p);
    if (c > maximum) return 0;
    uint64_t dist = int_distance_core(p, maximum - c);
    uint64_t sim = maximum - dist;
    return sim >= c ? sim : 0;
}

__device__ __forceinline__ double int_api_norm_distance(const IntPair& p, double c)
{
    uint64_t maximum = int_maximum(p);
    uint64_t cd = (uint64_t)ceil((double)maximum * c);
    double dist = (double)int_api_distance(p, cd);
    double norm = (maximum != 0) ? dist / (double)maximum : 0.0;
    return (norm <= c) ? norm : 1.0;
}

// NormSim_to_NormDist (common.hpp:50-53): std::min(1.0, 1 - c + 1e-5), NaN giving 1.0 as there
__device__ __forceinline__ double normsim_to_normdist(double c)
{
    const double w = 1.0 - c + 0.00001;
    return (w < 1.0) ? w : 1.0;
}

__device__ __forceinline__ double int_api_norm_similarity(const IntPair& p, double c)
{
    double nd = int_api_norm_distance(p, normsim_to_normdist(c));
    double ns = 1.0 - nd;
    return (ns >= c) ? ns : 0.0;
}

// Multi* semantics: the full (uncut) distance first, cut after.
__device__ __forceinline__ uint64_t multi_full_distance(const IntPair& p)
{
    switch (p.s->metric) {
    case RFA_INDEL:
    case RFA_RATIO: return indel_from_lcs(p);
    case RFA_LCSSEQ: return u64max(p.l1, p.l2) - p.raw;
    default: return p.raw; // Levenshtein (always the uniform recurrence here) / OSA
    }
}

// --- Jaro / Jaro-Winkler ---------------------------------------------------

// jaro_similarity(PM, P, T, cutoff) from the counts (Jaro_impl.hpp:32-71, 409-454).
//
// The reference checks two upper bounds of Sim against the cutoff before the
// formula: the length bound (min/P + min/T + 1) / 3 and the common-character
// bound (cc/P + cc/T + 1) / 3.  Each bound adds terms at least as large as the
// formula's (cc <= min(P, T), and (cc - t) / cc <= 1), in the same order, and
// rounded division and addition are monotone, so each computed bound is at
// least the computed Sim.  A failed bound therefore means Sim < cutoff, and the
// result is 0 either way: the formula and its final comparison decide alone.
// (1 x 1 strings reach their rule whenever cutoff <= 1, where the length bound
// is exactly 1.0.)  Lengths and counts are below 2^32 (strings are capped there),
// so they convert to double through 32 bits.
__device__ __forceinline__ double jaro_from_counts(uint64_t Pl, uint64_t Tl, uint64_t cc, uint64_t tr, bool eq11,
                                                   double cutoff)
{
    if (cutoff > 1.0) return 0.0;
    if (!Pl && !Tl) return 1.0;
    if (!Tl || !Pl) return 0.0;
    if (cutoff != cutoff) return 0.0; // a NaN cutoff fails the length bound
    if (Pl == 1 && Tl == 1) return eq11 ? 1.0 : 0.0;
    if (!cc) return 0.0;
    const uint32_t t = (uint32_t)tr / 2;
    const double dcc = (double)(uint32_t)cc;
    double Sim = 0;
    Sim += dcc / (double)(uint32_t)Pl;
    Sim += dcc / (double)(uint32_t)Tl;
    Sim += (dcc - (double)t) / dcc;
    Sim /= 3.0;
    return (Sim >= cutoff) ? Sim : 0;
}

struct FpPair {
    const ScoreSpec* s;
    uint64_t l1, l2;
    RawResult r;
};

__device__ __forceinline__ double fp_similarity(const FpPair& p, double c)
{
    const bool eq11 = p.r.v == 1; // 1x1 strings: one common char iff equal
    if (p.s->metric == RFA_JARO) return jaro_from_counts(p.l1, p.l2, p.r.v, p.r.tr, eq11, c);
    // JaroWinkler_impl.hpp:41-71
    double w = p.s->prefix_weight;
    double jaro_cutoff = c;
    if (jaro_cutoff > 0.7) {
        double prefix_sim = (double)p.r.prefix * w;
        if (prefix_sim >= 1.0)
            jaro_cutoff = 0.7;
        else
            jaro_cutoff = dmax(0.7, (prefix_sim - jaro_cutoff) / (prefix_sim - 1.0));
    }
    double Sim = jaro_from_counts(p.l1, p.l2, p.r.v, p.r.tr, eq11, jaro_cutoff);
    if (Sim > 0.7) {
        Sim += (double)p.r.prefix * w * (1.0 - Sim);
        Sim = dmin(Sim, 1.0);
    }
    return (Sim >= c) ? Sim : 0;
}

__device__ __forceinline__ double fp_distance(const FpPair& p, double c)
{
    double cs = (1.0 > c) ? 1.0 - c : 0;
    double dist = 1.0 - fp_similarity(p, cs);
    return (dist <= c) ? dist : 1.0;
}

__device__ __forceinline__ double fp_norm_distance(const FpPair& p, double c)
{
    double cd = ceil(1.0 * c);
    double dist = fp_distance(p, cd);
    double norm = dist / 1.0;
    return (norm <= c) ? norm : 1.0;
}

__device__ __forceinline__ double fp_norm_similarity(const FpPair& p, double c)
{
    double nd = fp_norm_distance(p, normsim_to_normdist(c));
    double ns = 1.0 - nd;
    return (ns >= c) ? ns : 0.0;
}

__device__ __forceinline__ double multi_fp_similarity(const FpPair& p, double c)
{
    const bool eq11 = p.r.v == 1;
    if (p.s->metric == RFA_JARO) return jaro_from_counts(p.l1, p.l2, p.r.v, p.r.tr, eq11, c);
    double s = jaro_from_counts(p.l1, p.l2, p.r.v, p.r.tr, eq11, dmin(0.7, c));
    if (s > 0.7) {
        s += (double)p.r.prefix * p.s->prefix_weight * (1.0 - s);
        s = dmin(s, 1.0);
    }
    if (s < c) s = 0.0;
    return s;
}

// The raw distance a fixed integer cutoff needs exactly (larger values only
// need to stay larger): distance kinds of Levenshtein (in units of its common
// weight), Indel and OSA; everything else needs the exact value.
__device__ __forceinline__ uint64_t band_from_cutoff(const ScoreSpec& s)
{
    if (s.kind != RFA_DISTANCE) return ~0ull;
    if (s.metric == RFA_LEVENSHTEIN) return s.ins ? ceil_div_u64(s.cutoff_u, s.ins) : ~0ull;
    if (s.metric == RFA_INDEL || s.metric == RFA_OSA) return s.cutoff_u;
    return ~0ull;
}

// --- results decided by the lengths -------------------------------------------

// The raw value of a pair of lengths m, n lies in [raw_range_lo, raw_range_hi]:
// a distance (Levenshtein, OSA) in [|m - n|, max(m, n)], an LCS length in
// [0, min(m, n)].
template <int ALGO>
__device__ __forceinline__ uint64_t raw_range_lo(uint64_t m, uint64_t n)
{
    if constexpr (ALGO == ALGO_LCS)
        return 0;
    else
        return m > n ? m - n : n - m;
}

template <int ALGO>
__device__ __forceinline__ uint64_t raw_range_hi(uint64_t m, uint64_t n)
{
    if constexpr (ALGO == ALGO_LCS)
        return m < n ? m : n;
    else
        return m > n ? m : n;
}


// --- dispatch ----------------------------------------------------------------

// Integer-valued kinds: distance / similarity of Levenshtein, Indel, LCSseq, OSA.
__device__ __forceinline__ uint64_t finalize_int(const ScoreSpec& s, uint64_t l1, uint64_t l2, uint64_t raw)
{
    IntPair p{&s, l1, l2, raw};
    if (s.semantics == RFA_MULTI) {
        // MultiLevenshtein's maximum() uses its weights (Levenshtein.hpp:401-404)
        const uint64_t maximum = int_maximum(p);
        const uint64_t d = multi_full_distance(p);
        if (s.kind == RFA_DISTANCE) return d <= s.cutoff_u ? d : s.cutoff_u + 1;
        if (s.metric == RFA_LCSSEQ) return raw >= s.cutoff_u ? raw : 0;
        const uint64_t sim = maximum - d;
        return sim >= s.cutoff_u ? sim : 0;
    }
    return s.kind == RFA_DISTANCE ? int_api_distance(p, s.cutoff_u) : int_api_similarity(p, s.cutoff_u);
}

// Jaro / Jaro-Winkler: the reference API value of one pair.
__device__ __forceinline__ double finalize_jaro(const ScoreSpec& s, uint64_t l1, uint64_t l2, RawResult r)
{
    FpPair p{&s, l1, l2, r};
    double v;
    if (s.semantics == RFA_MULTI) {
        if (s.kind == RFA_SIMILARITY)
            v = multi_fp_similarity(p, s.cutoff_d);
        else {
            double dist = 1.0 - multi_fp_similarity(p, 0.0);
            if (s.kind != RFA_DISTANCE) dist = dist <= 1.0 ? dist : 1.0;
            if (s.kind == RFA_NORMALIZED_SIMILARITY) {
                double ns = 1.0 - dist;
                v = ns >= s.cutoff_d ? ns : 0.0;
            }
            else
                v = dist <= s.cutoff_d ? dist : 1.0;
        }
    }
    else {
        switch (s.kind) {
        case RFA_DISTANCE: v = fp_distance(p, s.cutoff_d); break;
        case RFA_SIMILARITY: v = fp_similarity(p, s.cutoff_d); break;
        case RFA_NORMALIZED_DISTANCE: v = fp_norm_distance(p, s.cutoff_d); break;
        default: v = fp_norm_similarity(p, s.cutoff_d); break;
        }
    }
    return v;
}

// The integer-valued metrics (Levenshtein, Indel, LCSseq, OSA, ratio, the
// positional metrics): writes the value for one pair into out[idx] (uint64 or double).
// Normalized scores (and ratio) of Cached scorers whose maximum is max(l1, l2)
// or l1 + l2 (Levenshtein with unit weights, OSA, Indel, LCSseq, ratio) for a
// cutoff c' in [0, 1] in distance units: the chain of int_api_norm_distance /
// int_api_norm_similarity with 32-bit conversions and the cut decided before
// the division.  A cut distance d > cd = ceil(max * c') makes the reference
// divide cd + 1 (or more) by max; cd + 1 exceeds max * c' by nearly 1, so the
// quotient exceeds c' by far more than its rounding error and the distance
// becomes 1.0 (the similarity 0.0).  NormEpi holds the launch-level choices;
// the per-pair steps are straight-line selects.
struct NormEpi {
    bool lcs_raw; // the raw value is an LCS length (Indel, LCSseq, ratio), else a distance
    bool sum_max; // the maximum is l1 + l2 (Indel, ratio), else max(l1, l2)
    bool sim;     // normalized similarity (or ratio), else normalized distance
    bool ratio;   // times 100
    double cp;    // the cutoff in distance units
    double c_in;  // the similarity cutoff (sim)
};

// Whether the fast chain covers the launch; fills e.  (Also used on the host.)
__device__ __host__ inline bool norm_epi_setup(const ScoreSpec& s, NormEpi& e)
{
    if (s.semantics != RFA_CACHED) return false;
    e.ratio = s.metric == RFA_RATIO;
    if (!e.ratio && s.kind != RFA_NORMALIZED_DISTANCE && s.kind != RFA_NORMALIZED_SIMILARITY) return false;
    if (e.ratio && s.empty_zero) return false;
    e.lcs_raw = e.ratio || s.metric == RFA_INDEL || s.metric == RFA_LCSSEQ;
    e.sum_max = e.ratio || s.metric == RFA_INDEL;
    const bool lev_unit = s.metric == RFA_LEVENSHTEIN && s.ins == 1 && s.del == 1 && s.rep == 1;
    if (!(e.lcs_raw || lev_unit || s.metric == RFA_OSA)) return false;
    e.c_in = e.ratio ? s.cutoff_d / 100 : s.cutoff_d;
    e.sim = e.ratio || s.kind == RFA_NORMALIZED_SIMILARITY;
    // NormSim_to_NormDist (common.hpp:50-53)
    const double widened = 1.0 - e.c_in + 0.00001;
    e.cp = e.sim ? (widened < 1.0 ? widened : 1.0) : e.c_in;
    return e.cp >= 0.0 && e.cp <= 1.0;
}

__device__ __forceinline__ uint32_t norm_epi_max(const NormEpi& e, uint32_t l1, uint32_t l2)
{
    return e.sum_max ? l1 + l2 : max(l1, l2);
}

// the largest distance that is not cut
__device__ __forceinline__ uint32_t norm_epi_cut(const NormEpi& e, uint32_t M)
{
    return (uint32_t)ceil((double)M * e.cp);
}

// the exact distance: the raw value, or the maximum less one or two LCS lengths
__device__ __forceinline__ uint32_t norm_epi_dist(const NormEpi& e, uint32_t M, uint32_t raw)
{
    return e.lcs_raw ? M - (e.sum_max ? raw + raw : raw) : raw;
}

__device__ __forceinline__ double norm_epi_value(const NormEpi& e, uint32_t M, uint32_t d, uint32_t cd)
{
    double norm = 1.0;
    if (d <= cd) {
        norm = M ? (double)d / (double)M : 0.0;
        norm = (norm <= e.cp) ? norm : 1.0;
    }
    if (!e.sim) return norm;
    const double ns = 1.0 - norm;
    const double v = (ns >= e.c_in) ? ns : 0.0;
    return e.ratio ? v * 100 : v;
}

__device__ __forceinline__ bool norm_fast(const ScoreSpec& s, uint64_t l1, uint64_t l2, uint64_t raw, double& out)
{
    NormEpi e;
    if (!norm_epi_setup(s, e)) return false;
    const uint64_t M64 = e.sum_max ? l1 + l2 : u64max(l1, l2);
    if (M64 > 0xffffffffull) return false;
    const uint32_t M = (uint32_t)M64;
    out = norm_epi_value(e, M, norm_epi_dist(e, M, (uint32_t)raw), norm_epi_cut(e, M));
    return true;
}

// For an LCS-valued pair of lengths m, n (Indel, ratio, LCSseq; Cached
// semantics): a length T such that every LCS below T gives the same result,
// the cut value (an Indel distance M - 2 LCS above the integer cutoff c, an
// LCSseq distance max - LCS above c, a similarity below its cutoff), or 0 when
// no such bound applies.  The normalized kinds use the integer cutoff of
// norm_fast, where a cut distance gives 1.0 / 0.0.
__device__ __forceinline__ uint32_t lcs_cut_below(const ScoreSpec& s, uint64_t m, uint64_t n)
{
    if (s.semantics != RFA_CACHED) return 0;
    const bool ratio = s.metric == RFA_RATIO;
    if (!ratio && s.metric != RFA_INDEL && s.metric != RFA_LCSSEQ) return 0;
    if (ratio && s.empty_zero) return 0;
    const bool sum_max = s.metric != RFA_LCSSEQ;
    const uint64_t M = sum_max ? m + n : u64max(m, n);
    uint64_t c; // the largest distance that is not cut
    if (!ratio && s.kind == RFA_DISTANCE) {
        c = s.cutoff_u;
    }
    else if (!ratio && s.kind == RFA_SIMILARITY) {
        if (s.metric == RFA_LCSSEQ) return (uint32_t)u64min(s.cutoff_u, 0xffffffffull); // LCS < c gives 0
        if (s.cutoff_u > M) return 0;
        c = M - s.cutoff_u; // int_api_similarity: cut where d > M - c
    }
    else {
        NormEpi e;
        if (!norm_epi_setup(s, e) || M > 0xffffffffull) return 0;
        c = norm_epi_cut(e, (uint32_t)M);
    }
    if (M <= c) return 0;
    // cut where the distance exceeds c: M - 2 LCS > c (Indel), max - LCS > c (LCSseq)
    return (uint32_t)(sum_max ? (M - c + 1) / 2 : M - c);
}

__device__ __forceinline__ uint64_t dbits(double d) { return __builtin_bit_cast(uint64_t, d); }

__device__ __forceinline__ uint64_t finalize_intmetric(const ScoreSpec& s, uint64_t l1, uint64_t l2, RawResult r)
{
    IntPair p{&s, l1, l2, r.v};
    if (s.metric == RFA_RATIO && s.empty_zero && (l1 == 0 || l2 == 0)) {
        return dbits(0.0); // CachedQRatio::similarity
    }
    {
        double v;
        if (norm_fast(s, l1, l2, r.v, v)) return dbits(v);
    }
    if (s.metric == RFA_RATIO) {
        // ratio = Indel normalized similarity * 100 (Cached: with the +1e-5 widening)
        ScoreSpec t = s;
        t.metric = RFA_INDEL;
        IntPair q{&t, l1, l2, r.v};
        if (s.semantics == RFA_MULTI) {
            uint64_t maximum = l1 + l2;
            double norm = maximum ? (double)indel_from_lcs(q) / (double)maximum : 0.0;
            norm = norm <= 1.0 ? norm : 1.0;
            double ns = 1.0 - norm;
            double c = s.cutoff_d / 100.0;
            return dbits((ns >= c ? ns : 0.0) * 100.0);
        }
        return dbits(int_api_norm_similarity(q, s.cutoff_d / 100) * 100);
    }
    if (s.kind == RFA_DISTANCE || s.kind == RFA_SIMILARITY) return finalize_int(s, l1, l2, r.v);
    if (s.semantics == RFA_MULTI) {
        uint64_t maximum = int_maximum(p);
        uint64_t d = multi_full_distance(p);
        switch (s.kind) {
        case RFA_NORMALIZED_DISTANCE: {
            double norm = maximum ? (double)d / (double)maximum : 0.0;
            return dbits(norm <= s.cutoff_d ? norm : 1.0);
        }
        default: {
            double norm = maximum ? (double)d / (double)maximum : 0.0;
            norm = norm <= 1.0 ? norm : 1.0;
            double ns = 1.0 - norm;
            return dbits(ns >= s.cutoff_d ? ns : 0.0);
        }
        }
    }
    if (s.kind == RFA_NORMALIZED_DISTANCE) return dbits(int_api_norm_distance(p, s.cutoff_d));
    return dbits(int_api_norm_similarity(p, s.cutoff_d));
}

// The reference API value of one pair as its 64 result bits (uint64 or the
// double's bit pattern).
__device__ __forceinline__ uint64_t finalize_bits(const ScoreSpec& s, uint64_t l1, uint64_t l2, RawResult r)
{
    if (s.metric == RFA_JARO || s.metric == RFA_JARO_WINKLER) return dbits(finalize_jaro(s, l1, l2, r));
    return finalize_intmetric(s, l1, l2, r);
}

// Writes the value into out[idx] (8-byte elements: uint64 or double).
__device__ __forceinline__ void finalize(const ScoreSpec& s, uint64_t l1, uint64_t l2, RawResult r, void* out,
                                         uint64_t idx)
{
    ((uint64_t*)out)[idx] = finalize_bits(s, l1, l2, r);
}

// finalize_bits for a kernel whose recurrence fixes the metric family: the Jaro
// epilogue only in Jaro kernels, the integer-metric epilogue everywhere else.
template <int ALGO>
__device__ __forceinline__ uint64_t finalize_bits_algo(const ScoreSpec& s, uint64_t l1, uint64_t l2, RawResult r)
{
    if constexpr (ALGO == ALGO_JARO)
        return dbits(finalize_jaro(s, l1, l2, r));
    else
        return finalize_intmetric(s, l1, l2, r);
}

} // namespace rfa
