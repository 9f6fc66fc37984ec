// rfa_types.hpp -- types shared by host dispatch and kernels.
#pragma once
#include <stdint.h>

#include "../../include/rapidfuzz_amd.h"

namespace rfa {

// What to compute for every pair of a launch (passed by value as a kernel argument).
struct ScoreSpec {
    int32_t metric;    // rfa_metric
    int32_t kind;      // rfa_kind
    int32_t semantics; // rfa_semantics
    int32_t algo;      // Algo: the bit-parallel recurrence that produces the raw value
    uint64_t cutoff_u;
    double cutoff_d;
    uint64_t ins, del, rep;
    double prefix_weight;
    uint64_t band_k; // exact answer needed only up to band_k (used to skip pairs by length)
    int32_t empty_zero; // QRatio: 0 when either string is empty (fuzz_impl.hpp:912-922)
    int32_t reserved;
};

// Raw recurrences.
enum Algo : int32_t {
    ALGO_LEV = 0,  // Hyyro 2003 Levenshtein (Levenshtein_impl.hpp:259-322, 641-850)
    ALGO_LCS = 1,  // Hyyro 2004 LCS (LCSseq_impl.hpp:205-333)
    ALGO_OSA = 2,  // Hyyro 2003 with transpositions (OSA_impl.hpp:34-74, 176-249)
    ALGO_JARO = 3, // greedy window flagging + transposition walk (Jaro_impl.hpp:94-315)
    ALGO_MIXED = 4, // per-item metric (Levenshtein / Indel / LCSseq / OSA), chosen at run time
    ALGO_POSIT = 5, // positional comparison: Hamming (Hamming_impl.hpp:23-35), common prefix / suffix
                    // (Prefix_impl.hpp:23-27, Postfix_impl.hpp:23-27)
    ALGO_PARTIAL = 6, // LCS of the shorter string against every window of the longer one (fuzz_impl.hpp:57-279)
    ALGO_TSET = 7,    // masked LCS of two token-set differences (fuzz_impl.hpp:370-426)
    ALGO_WLEV = 8,    // Wagner-Fischer with general weights (Levenshtein_impl.hpp:85-111)
    ALGO_DL = 9       // Damerau-Levenshtein, Zhao and Sahni (DamerauLevenshtein_impl.hpp:36-103)
};

// A packed string set on the device.
template <typename CharT>
struct DevStrings {
    const CharT* chars;
    const uint64_t* off; // count + 1
    uint64_t count;
};

} // namespace rfa
