/*
 * rf_oracle.c -- CPU restatement of rapidfuzz-cpp scoring (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this file's library, and only as a checker.  The product path never does.
 *
 * Raw values are naive dynamic programs restated from the reference's own
 * oracle directory; the API layer restates the result semantics of the
 * reference's CRTP bases step by step.  Each function names the reference
 * lines it follows.
 */
#include "rf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint64_t u64min(uint64_t a, uint64_t b) { return a < b ? a : b; }
static uint64_t u64max(uint64_t a, uint64_t b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }
/* std::max(a, b): a unless a < b */
static double std_max(double a, double b) { return (a < b) ? b : a; }

/* rapidfuzz/details/intrinsics.hpp ceil_div */
static uint64_t ceil_div(uint64_t a, uint64_t b) { return a / b + (uint64_t)(a % b != 0); }

/* ------------------------------------------------------------------------ */
/* raw values                                                               */
/* ------------------------------------------------------------------------ */

/* Wagner-Fischer with weights: rapidfuzz_reference/Levenshtein.hpp:31-56
 * (column-by-column with one rolling column instead of the full matrix). */
uint64_t orc_levenshtein_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2,
                             uint64_t ins, uint64_t del, uint64_t rep)
{
    uint64_t* col = (uint64_t*)malloc((l1 + 1) * sizeof(uint64_t));
    for (size_t i = 0; i <= l1; ++i) col[i] = (uint64_t)i * del;
    for (size_t j = 0; j < l2; ++j) {
        uint64_t diag = col[0];
        col[0] = (uint64_t)(j + 1) * ins;
        for (size_t i = 0; i < l1; ++i) {
            uint64_t cost = (s1[i] == s2[j]) ? 0 : rep;
            uint64_t up = col[i + 1]; /* D[i+1][j] */
            uint64_t v = up + ins;    /* insertion of s2[j] */
            uint64_t a = col[i] + del; /* deletion of s1[i] (D[i][j+1]) */
            uint64_t b = diag + cost;
            if (a < v) v = a;
            if (b < v) v = b;
            diag = up;
            col[i + 1] = v;
        }
    }
    uint64_t d = col[l1];
    free(col);
    return d;
}

/* rapidfuzz_reference/LCSseq.hpp:9-17: LCS = (len1 + len2 - Indel) / 2 with
 * Indel = Levenshtein with weights {1, 1, 2} (rapidfuzz_reference/Indel.hpp:11-16). */
uint64_t orc_lcs_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    uint64_t indel = orc_levenshtein_raw(s1, l1, s2, l2, 1, 1, 2);
    return ((uint64_t)l1 + (uint64_t)l2 - indel) / 2;
}

/* rapidfuzz_reference/OSA.hpp:13-43 (rolling three rows over s2). */
uint64_t orc_osa_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    size_t w = l2 + 1;
    uint64_t* m = (uint64_t*)malloc(3 * w * sizeof(uint64_t));
    uint64_t *prev2 = m, *prev = m + w, *cur = m + 2 * w;
    for (size_t j = 0; j <= l2; ++j) prev[j] = j;
    for (size_t i = 0; i < l1; ++i) {
        cur[0] = i + 1;
        for (size_t j = 0; j < l2; ++j) {
            uint64_t cost = (s1[i] == s2[j]) ? 0 : 1;
            uint64_t v = prev[j + 1] + 1;
            if (cur[j] + 1 < v) v = cur[j] + 1;
            if (prev[j] + cost < v) v = prev[j] + cost;
            if (i > 0 && j > 0 && s1[i] == s2[j - 1] && s1[i - 1] == s2[j]) {
                if (prev2[j - 1] + cost < v) v = prev2[j - 1] + cost;
            }
            cur[j + 1] = v;
        }
        uint64_t* t = prev2;
        prev2 = prev;
        prev = cur;
        cur = t;
    }
    uint64_t d = prev[l2];
    free(m);
    return d;
}

/* Damerau-Levenshtein after Zhao and Sahni, as damerau_levenshtein_distance_zhao
 * runs it (DamerauLevenshtein_impl.hpp:36-103): rows over s1, one row of s2
 * positions at a time, R / R1 the current and previous rows (R holds the row
 * before R1 until overwritten), FR the value saved at the last match of each s2
 * position, T the value saved at the last match of the current row.  The
 * reference's last_row_id map (s1 character -> last row it occurred in, read at
 * s2[j]) is kept here per s2 position: lastk[j] = last row i with s1[i] == s2[j],
 * written when that row finishes, which is the value the map returns.  maxVal is
 * max(l1, l2) + 1 as there. */
uint64_t orc_damerau_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    const int64_t maxv = (int64_t)(l1 > l2 ? l1 : l2) + 1;
    const size_t size = l2 + 2;
    int64_t* buf = (int64_t*)malloc(3 * size * sizeof(int64_t));
    int64_t* lastk = (int64_t*)malloc((l2 + 1) * sizeof(int64_t));
    int64_t *FRa = buf, *R1a = buf + size, *Ra = buf + 2 * size;
    for (size_t j = 0; j < size; ++j) FRa[j] = R1a[j] = maxv;
    Ra[0] = maxv;
    for (size_t j = 1; j < size; ++j) Ra[j] = (int64_t)j - 1;
    for (size_t j = 0; j <= l2; ++j) lastk[j] = -1;
    int64_t* R = Ra + 1; /* R[-1] is the sentinel column */
    int64_t* R1 = R1a + 1;
    int64_t* FR = FRa + 1;
    for (int64_t i = 1; i <= (int64_t)l1; ++i) {
        int64_t* t = R;
        R = R1;
        R1 = t;
        int64_t last_col_id = -1;
        int64_t last_i2l1 = R[0];
        R[0] = i;
        int64_t T = maxv;
        const uint32_t a = s1[i - 1];
        for (int64_t j = 1; j <= (int64_t)l2; ++j) {
            const uint32_t b = s2[j - 1];
            int64_t diag = R1[j - 1] + (a != b);
            int64_t left = R[j - 1] + 1;
            int64_t up = R1[j] + 1;
            int64_t temp = diag < left ? diag : left;
            temp = up < temp ? up : temp;
            if (a == b) {
                last_col_id = j;
                FR[j] = R1[j - 2];
                T = last_i2l1;
            }
            else {
                int64_t k = lastk[j];
                int64_t l = last_col_id;
                if (j - l == 1) {
                    int64_t tr = FR[j] + (i - k);
                    temp = tr < temp ? tr : temp;
                }
                else if (i - k == 1) {
                    int64_t tr = T + (j - l);
                    temp = tr < temp ? tr : temp;
                }
            }
            last_i2l1 = R[j];
            R[j] = temp;
        }
        for (size_t j = 1; j <= l2; ++j)
            if (s2[j - 1] == a) lastk[j] = i;
    }
    uint64_t d = (uint64_t)R[l2];
    free(buf);
    free(lastk);
    return d;
}

/* rapidfuzz_reference/Jaro.hpp:22-56: greedy matching window and
 * transposition walk.  Writes the common-character count and the number of
 * mismatched flagged pairs (before the halving of the Jaro formula). */
void orc_jaro_raw(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl, uint64_t* common,
                  uint64_t* transpositions)
{
    *common = 0;
    *transpositions = 0;
    if (!Pl || !Tl) return;
    unsigned char* Pf = (unsigned char*)calloc(Pl + 1, 1);
    unsigned char* Tf = (unsigned char*)calloc(Tl + 1, 1);
    size_t Bound = (Pl > Tl ? Pl : Tl) / 2;
    if (Bound > 0) Bound--;
    uint64_t cc = 0;
    for (size_t i = 0; i < Tl; i++) {
        size_t lowlim = (i >= Bound) ? i - Bound : 0;
        size_t hilim = (i + Bound <= Pl - 1) ? (i + Bound) : Pl - 1;
        for (size_t j = lowlim; j <= hilim; j++) {
            if (!Pf[j] && P[j] == T[i]) {
                Tf[i] = 1;
                Pf[j] = 1;
                cc++;
                break;
            }
        }
    }
    uint64_t tr = 0;
    size_t k = 0;
    for (size_t i = 0; i < Tl; i++) {
        if (Tf[i]) {
            size_t j = k;
            for (; j < Pl; j++) {
                if (Pf[j]) {
                    k = j + 1;
                    break;
                }
            }
            if (T[i] != P[j]) tr++;
        }
    }
    free(Pf);
    free(Tf);
    *common = cc;
    *transpositions = tr;
}

/* rapidfuzz/distance/Jaro_impl.hpp:32-42 (operation order kept). */
static double jaro_formula(size_t Pl, size_t Tl, uint64_t cc, uint64_t tr)
{
    tr /= 2;
    double Sim = 0;
    Sim += (double)cc / (double)Pl;
    Sim += (double)cc / (double)Tl;
    Sim += ((double)cc - (double)tr) / (double)cc;
    return Sim / 3.0;
}

/* Library-level Jaro with cutoff: rapidfuzz/distance/Jaro_impl.hpp:409-454
 * with the filters of :47-71.  The filters are kept literally so the double
 * results match bit for bit. */
double orc_jaro_similarity(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl, double cutoff)
{
    if (cutoff > 1.0) return 0.0;
    if (!Pl && !Tl) return 1.0;
    /* jaro_length_filter */
    if (!Tl || !Pl) return 0.0;
    {
        double min_len = (double)(Pl < Tl ? Pl : Tl);
        double Sim = min_len / (double)Pl + min_len / (double)Tl + 1.0;
        Sim /= 3.0;
        if (!(Sim >= cutoff)) return 0.0;
    }
    if (Pl == 1 && Tl == 1) return (double)(P[0] == T[0]);
    uint64_t cc, tr;
    orc_jaro_raw(P, Pl, T, Tl, &cc, &tr);
    /* jaro_common_char_filter */
    if (!cc) return 0.0;
    {
        double Sim = 0;
        Sim += (double)cc / (double)Pl;
        Sim += (double)cc / (double)Tl;
        Sim += 1.0;
        Sim /= 3.0;
        if (!(Sim >= cutoff)) return 0.0;
    }
    double Sim = jaro_formula(Pl, Tl, cc, tr);
    return (Sim >= cutoff) ? Sim : 0;
}

static size_t common_prefix4(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl)
{
    size_t max_prefix = Pl < Tl ? Pl : Tl;
    if (max_prefix > 4) max_prefix = 4;
    size_t prefix = 0;
    for (; prefix < max_prefix; ++prefix)
        if (T[prefix] != P[prefix]) break;
    return prefix;
}

/* rapidfuzz/distance/JaroWinkler_impl.hpp:41-71 (cached path). */
double orc_jaro_winkler_similarity(const uint32_t* P, size_t Pl, const uint32_t* T, size_t Tl,
                                   double prefix_weight, double cutoff)
{
    size_t prefix = common_prefix4(P, Pl, T, Tl);
    double jaro_cutoff = cutoff;
    if (jaro_cutoff > 0.7) {
        double prefix_sim = (double)prefix * prefix_weight;
        if (prefix_sim >= 1.0)
            jaro_cutoff = 0.7;
        else
            jaro_cutoff = dmax(0.7, (prefix_sim - jaro_cutoff) / (prefix_sim - 1.0));
    }
    double Sim = orc_jaro_similarity(P, Pl, T, Tl, jaro_cutoff);
    if (Sim > 0.7) {
        Sim += (double)prefix * prefix_weight * (1.0 - Sim);
        Sim = dmin(Sim, 1.0);
    }
    return (Sim >= cutoff) ? Sim : 0;
}

/* ------------------------------------------------------------------------ */
/* Cached* API layer                                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    int metric;
    const uint32_t *s1, *s2;
    size_t l1, l2;
    uint64_t ins, del, rep;
    double pw;
} Pair;

/* rapidfuzz/distance/Levenshtein_impl.hpp:118-128 */
static uint64_t lev_maximum(uint64_t l1, uint64_t l2, uint64_t ins, uint64_t del, uint64_t rep)
{
    uint64_t m = l1 * del + l2 * ins;
    if (l1 >= l2)
        m = u64min(m, l2 * rep + (l1 - l2) * del);
    else
        m = u64min(m, l1 * rep + (l2 - l1) * ins);
    return m;
}

/* Hamming with padding (Hamming_impl.hpp:23-35; the naive restatement in
 * rapidfuzz_reference/Hamming.hpp:10-23 without its equal-length requirement):
 * max(l1, l2) minus the equal positions of the common length. */
uint64_t orc_hamming_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    size_t n = l1 < l2 ? l1 : l2;
    uint64_t dist = l1 > l2 ? l1 : l2;
    for (size_t i = 0; i < n; ++i) dist -= (s1[i] == s2[i]);
    return dist;
}

/* Prefix::_similarity (Prefix_impl.hpp:23-27): length of the common prefix. */
uint64_t orc_prefix_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    size_t i = 0;
    while (i < l1 && i < l2 && s1[i] == s2[i]) ++i;
    return i;
}

/* Postfix::_similarity (Postfix_impl.hpp:23-27): length of the common suffix. */
uint64_t orc_postfix_raw(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    size_t i = 0;
    while (i < l1 && i < l2 && s1[l1 - 1 - i] == s2[l2 - 1 - i]) ++i;
    return i;
}

/* CachedPrefix / CachedPostfix::_similarity: the length when it reaches the
 * cutoff, else 0 (Prefix.hpp:84-88, Postfix.hpp:84-88). */
static uint64_t affix_cached_similarity(const Pair* p, uint64_t c)
{
    uint64_t sim = p->metric == ORC_PREFIX ? orc_prefix_raw(p->s1, p->l1, p->s2, p->l2)
                                           : orc_postfix_raw(p->s1, p->l1, p->s2, p->l2);
    return sim >= c ? sim : 0;
}

/* CachedHamming::_distance (Hamming.hpp:153-157, pad = true). */
static uint64_t hamming_cached_distance(const Pair* p, uint64_t c)
{
    uint64_t d = orc_hamming_raw(p->s1, p->l1, p->s2, p->l2);
    return d <= c ? d : c + 1;
}

/* CachedLCSseq::_similarity -> lcs_seq_similarity (LCSseq_impl.hpp:375-403):
 * the exact LCS when it reaches the cutoff, else 0. */
static uint64_t lcs_cached_similarity(const Pair* p, uint64_t c)
{
    if (c > p->l1 || c > p->l2) return 0;
    uint64_t lcs = orc_lcs_bits(p->s1, p->l1, p->s2, p->l2); /* equals orc_lcs_raw's DP */
    return lcs >= c ? lcs : 0;
}

/* detail::indel_distance(PM, s1, s2, cutoff) (Indel_impl.hpp:15-24) and
 * CachedIndel::_distance (Indel.hpp:171-179). */
static uint64_t indel_cached_distance(const Pair* p, uint64_t c)
{
    uint64_t maximum = (uint64_t)p->l1 + (uint64_t)p->l2;
    uint64_t lcs_cutoff = (maximum / 2 >= c) ? maximum / 2 - c : 0;
    uint64_t lcs_sim = lcs_cached_similarity(p, lcs_cutoff);
    uint64_t dist = maximum - 2 * lcs_sim;
    return (dist <= c) ? dist : c + 1;
}

/* uniform_levenshtein_distance result contract (Levenshtein_impl.hpp:852-905):
 * exact distance when <= the clamped cutoff, else cutoff + 1. */
static uint64_t lev_uniform_cached(const Pair* p, uint64_t c)
{
    c = u64min(c, u64max(p->l1, p->l2));
    uint64_t d = orc_levenshtein_raw(p->s1, p->l1, p->s2, p->l2, 1, 1, 1);
    return d <= c ? d : c + 1;
}

/* CachedLevenshtein::_distance (Levenshtein.hpp:445-477). */
static uint64_t lev_cached_distance(const Pair* p, uint64_t c)
{
    if (p->ins == p->del) {
        if (p->ins == 0) return 0;
        if (p->ins == p->rep) {
            uint64_t nc = ceil_div(c, p->ins);
            uint64_t dist = lev_uniform_cached(p, nc) * p->ins;
            return (dist <= c) ? dist : c + 1;
        }
        else if (p->rep >= p->ins + p->del) {
            uint64_t nc = ceil_div(c, p->ins);
            uint64_t dist = indel_cached_distance(p, nc) * p->ins;
            return (dist <= c) ? dist : c + 1;
        }
    }
    /* generalized_levenshtein_distance (Levenshtein_impl.hpp:141-152) */
    uint64_t min_edits = p->l1 > p->l2 ? (p->l1 - p->l2) * p->del : (p->l2 - p->l1) * p->ins;
    if (min_edits > c) return c + 1;
    uint64_t d = orc_levenshtein_raw(p->s1, p->l1, p->s2, p->l2, p->ins, p->del, p->rep);
    return d <= c ? d : c + 1;
}

/* ---- editops --------------------------------------------------------------
 * levenshtein_editops (Levenshtein_impl.hpp:1229-1257) and lcs_seq_editops
 * (LCSseq_impl.hpp:528-534) strip the common prefix and suffix, record the
 * bit-parallel matrix of the remaining strings and walk it back
 * (recover_alignment, Levenshtein_impl.hpp:959-1015, LCSseq_impl.hpp:440-506).
 * The bits are restated here from plain DP tables: for s1 = a (columns) and
 * s2 = b (rows), VP / VN of row r and column c mark D[c][r] - D[c-1][r] = +1 / -1
 * (D the Levenshtein table over prefixes), and the LCS state S has the bit set
 * where L[c][r] == L[c-1][r] (no LCS step at column c).  Levenshtein's walk is
 * the reference's only while the stripped s1 has at most 64 units (longer
 * ones record banded or Hirschberg matrices); LCS's holds for every length. */
enum { ED_REPLACE = 1, ED_INSERT = 2, ED_DELETE = 3 };

/* levenshtein_align for a stripped first string of more than 64 units below the
 * Hirschberg split (Levenshtein_impl.hpp:1023-1045, 1185-1200): the block recurrence
 * with RecordMatrix and max = the longer length (Levenshtein_impl.hpp:641-850),
 * Ukkonen's band of blocks moving per row, then recover_alignment
 * (Levenshtein_impl.hpp:959-1015) over the ShiftedBitMatrix rows
 * (details/Matrix.hpp:121-180: bits left of a row's first block read clear; a row's
 * words past the band keep the initial VP = ~0, VN = 0).  Returns the operation
 * count, or SIZE_MAX if the band vanishes. */
static size_t lev_block_editops(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t pre,
                                int32_t* type, uint64_t* src_pos, uint64_t* dest_pos)
{
    const size_t words = (la + 63) / 64;
    size_t max = la > lb ? la : lb;
    const size_t full_band = la < 2 * max + 1 ? la : 2 * max + 1;
    size_t fbw = full_band / 64 + 2;
    if (fbw > words) fbw = words;
    uint64_t* VPm = (uint64_t*)malloc(lb * fbw * sizeof(uint64_t));
    uint64_t* VNm = (uint64_t*)malloc(lb * fbw * sizeof(uint64_t));
    size_t* off = (size_t*)malloc(lb * sizeof(size_t));
    uint64_t* VP = (uint64_t*)malloc(words * sizeof(uint64_t));
    uint64_t* VN = (uint64_t*)malloc(words * sizeof(uint64_t));
    ptrdiff_t* scores = (ptrdiff_t*)malloc(words * sizeof(ptrdiff_t));
    for (size_t i = 0; i < lb * fbw; ++i) VPm[i] = ~0ull, VNm[i] = 0;
    for (size_t w = 0; w < words; ++w) VP[w] = ~0ull, VN[w] = 0, scores[w] = (ptrdiff_t)((w + 1) * 64);
    scores[words - 1] = (ptrdiff_t)la;
    const uint64_t Last = 1ull << ((la - 1) % 64);
    ptrdiff_t first = 0, last;
    {
        size_t t = (max + la - lb) / 2;
        if (max < t) t = max;
        size_t nb = (t + 1 + 63) / 64;
        if (nb > words) nb = words;
        last = (ptrdiff_t)nb - 1;
    }
    size_t result = 0;
    int failed = 0;
    for (size_t row = 0; row < lb && !failed; ++row) {
        uint64_t HP_carry = 1, HN_carry = 0;
        off[row] = (size_t)first * 64;
#define ADVANCE(word, delta)                                                                                  \
    do {                                                                                                     \
        uint64_t PMj = 0;                                                                                    \
        for (size_t i = 0; i < 64 && (word) * 64 + i < la; ++i)                                              \
            if (a[(word) * 64 + i] == b[row]) PMj |= 1ull << i;                                              \
        const uint64_t X = PMj | HN_carry;                                                                   \
        const uint64_t D0 = (((X & VP[word]) + VP[word]) ^ VP[word]) | X | VN[word];                         \
        uint64_t HP = VN[word] | ~(D0 | VP[word]);                                                           \
        uint64_t HN = D0 & VP[word];                                                                         \
        const uint64_t HPt = HP_carry, HNt = HN_carry;                                                       \
        if ((size_t)(word) < words - 1) {                                                                    \
            HP_carry = HP >> 63;                                                                             \
            HN_carry = HN >> 63;                                                                             \
        }                                                                                                    \
        else {                                                                                               \
            HP_carry = (HP & Last) != 0;                                                                     \
            HN_carry = (HN & Last) != 0;                                                                     \
        }                                                                                                    \
        HP = (HP << 1) | HPt;                                                                                \
        HN = (HN << 1) | HNt;                                                                                \
        VP[word] = HN | ~(D0 | HP);                                                                          \
        VN[word] = HP & D0;                                                                                  \
        VPm[row * fbw + (size_t)((word) - first)] = VP[word];                                                \
        VNm[row * fbw + (size_t)((word) - first)] = VN[word];                                                \
        (delta) = (ptrdiff_t)HP_carry - (ptrdiff_t)HN_carry;                                                 \
    } while (0)
        for (ptrdiff_t w = first; w <= last; ++w) {
            ptrdiff_t d;
            ADVANCE(w, d);
            scores[w] += d;
        }
        {
            const ptrdiff_t r1 = (ptrdiff_t)lb - (ptrdiff_t)row - 1;
            const ptrdiff_t r2 = (ptrdiff_t)la - ((ptrdiff_t)((1 + last) * 64 - 1) - 1);
            const ptrdiff_t m2 = scores[last] + (r1 > r2 ? r1 : r2);
            if (m2 < (ptrdiff_t)max) max = (size_t)m2;
        }
#define ROW_NUM(w) ((size_t)(w) + 1 == words ? la - 1 : ((size_t)(w) + 1) * 64 - 1)
        if ((size_t)last + 1 < words) {
            const ptrdiff_t cond = (ptrdiff_t)(max + 128 + row + la) - (ptrdiff_t)(scores[last] + 2 + lb);
            if ((ptrdiff_t)ROW_NUM(last) < cond) {
                ++last;
                VP[last] = ~0ull;
                VN[last] = 0;
                const size_t chars = (size_t)last + 1 == words ? ((la - 1) % 64 + 1) : 64;
                scores[last] = scores[last - 1] + (ptrdiff_t)chars - (ptrdiff_t)HP_carry + (ptrdiff_t)HN_carry;
                ptrdiff_t d;
                ADVANCE(last, d);
                scores[last] += d;
            }
        }
        for (; last >= first; --last) {
            const int c1 = scores[last] < (ptrdiff_t)(max + 64);
            const ptrdiff_t cond = (ptrdiff_t)(max + 128 + row + la + 1) - (ptrdiff_t)(scores[last] + 2 + lb);
            const int c2 = (ptrdiff_t)ROW_NUM(last) <= cond;
            if (c1 && c2) break;
        }
        for (; first <= last; ++first) {
            const int c1 = scores[first] < (ptrdiff_t)(max + 64);
            const ptrdiff_t cond = (ptrdiff_t)(scores[first] + la + row) - (ptrdiff_t)(max + lb);
            const int c2 = (ptrdiff_t)ROW_NUM(first) >= cond;
            if (c1 && c2) break;
        }
        if (last < first) failed = 1;
#undef ROW_NUM
#undef ADVANCE
    }
    if (!failed) {
        size_t dist = (size_t)scores[words - 1];
        if (dist > max) failed = 1;
        result = dist;
        size_t col = la, row = lb;
#define TEST_BIT(M, r, c)                                                                                     \
    ((c) < off[r] ? 0 : (int)(((M)[(r) * fbw + ((c) - off[r]) / 64] >> (((c) - off[r]) % 64)) & 1))
        while (!failed && row && col) {
            if (TEST_BIT(VPm, row - 1, col - 1)) {
                --col;
                --dist;
                type[dist] = ED_DELETE, src_pos[dist] = col + pre, dest_pos[dist] = row + pre;
            }
            else {
                --row;
                if (row && TEST_BIT(VNm, row - 1, col - 1)) {
                    --dist;
                    type[dist] = ED_INSERT, src_pos[dist] = col + pre, dest_pos[dist] = row + pre;
                }
                else {
                    --col;
                    if (a[col] != b[row]) {
                        --dist;
                        type[dist] = ED_REPLACE, src_pos[dist] = col + pre, dest_pos[dist] = row + pre;
                    }
                }
            }
        }
#undef TEST_BIT
        while (!failed && col) {
            --col;
            --dist;
            type[dist] = ED_DELETE, src_pos[dist] = col + pre, dest_pos[dist] = row + pre;
        }
        while (!failed && row) {
            --row;
            --dist;
            type[dist] = ED_INSERT, src_pos[dist] = col + pre, dest_pos[dist] = row + pre;
        }
    }
    free(VPm);
    free(VNm);
    free(off);
    free(VP);
    free(VN);
    free(scores);
    return failed ? (size_t)-1 : result;
}

/* 1: Levenshtein pairs past 64 stripped units follow the reference's block band (the
 * default); 0: the full-matrix walk, kept to show where the two differ */
static int g_lev_block = 1;
void orc_set_lev_block(int on) { g_lev_block = on; }

/* The walk of the full DP table of a stripped pair (Levenshtein or LCS): the bits the
 * reference records read from the table's differences; operations at positions
 * (col + src_off, row + dst_off). */
static size_t ed_table_walk(int metric, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t src_off,
                            size_t dst_off, int32_t* type, uint64_t* src_pos, uint64_t* dest_pos)
{
    const size_t W = la + 1;
    uint64_t* T = (uint64_t*)malloc((la + 1) * (lb + 1) * sizeof(uint64_t)); /* T[r * W + c] */
    const int lev = metric == ORC_LEVENSHTEIN;
    for (size_t c = 0; c <= la; ++c) T[c] = lev ? c : 0;
    for (size_t r = 1; r <= lb; ++r) {
        T[r * W] = lev ? r : 0;
        for (size_t c = 1; c <= la; ++c) {
            uint64_t v;
            if (lev) {
                v = T[(r - 1) * W + c - 1] + (a[c - 1] != b[r - 1]);
                if (T[(r - 1) * W + c] + 1 < v) v = T[(r - 1) * W + c] + 1;
                if (T[r * W + c - 1] + 1 < v) v = T[r * W + c - 1] + 1;
            }
            else {
                if (a[c - 1] == b[r - 1])
                    v = T[(r - 1) * W + c - 1] + 1;
                else
                    v = T[(r - 1) * W + c] > T[r * W + c - 1] ? T[(r - 1) * W + c] : T[r * W + c - 1];
            }
            T[r * W + c] = v;
        }
    }
    size_t dist = lev ? (size_t)T[lb * W + la] : la + lb - 2 * (size_t)T[lb * W + la];
    const size_t total = dist;
    size_t col = la, row = lb;
#define ED_PUT(t)                                                                                                    \
    do {                                                                                                         \
        --dist;                                                                                                  \
        type[dist] = (t);                                                                                        \
        src_pos[dist] = col + src_off;                                                                           \
        dest_pos[dist] = row + dst_off;                                                                           \
    } while (0)
    if (lev) {
        while (row && col) {
            /* VP bit (row - 1, col - 1): D[col][row] - D[col-1][row] == +1 */
            if (T[row * W + col] == T[row * W + col - 1] + 1) {
                --col;
                ED_PUT(ED_DELETE);
            }
            else {
                --row;
                /* VN bit (row - 1, col - 1) of the row above: D[col][row] - D[col-1][row] == -1 */
                if (row && T[row * W + col] + 1 == T[row * W + col - 1]) {
                    ED_PUT(ED_INSERT);
                }
                else {
                    --col;
                    if (a[col] != b[row]) ED_PUT(ED_REPLACE);
                }
            }
        }
    }
    else {
        while (row && col) {
            /* S bit (row - 1, col - 1): no LCS step at column col */
            if (T[row * W + col] == T[row * W + col - 1]) {
                --col;
                ED_PUT(ED_DELETE);
            }
            else {
                --row;
                /* S bit (row - 1, col - 1) clear: an LCS step at column col already by this row */
                if (row && T[row * W + col] != T[row * W + col - 1]) {
                    ED_PUT(ED_INSERT);
                }
                else {
                    --col;
                }
            }
        }
    }
    while (col) {
        --col;
        ED_PUT(ED_DELETE);
    }
    while (row) {
        --row;
        ED_PUT(ED_INSERT);
    }
#undef ED_PUT
    free(T);
    return total;
}

/* Levenshtein rows: D[i] = distance(a[0, i), b) for i = 0..la (the DP's last row) */
static void lev_last_row(const uint32_t* a, size_t la, const uint32_t* b, size_t lb, uint64_t* D)
{
    for (size_t i = 0; i <= la; ++i) D[i] = i;
    for (size_t r = 1; r <= lb; ++r) {
        uint64_t diag = D[0];
        D[0] = r;
        for (size_t i = 1; i <= la; ++i) {
            const uint64_t up = D[i];
            uint64_t v = diag + (a[i - 1] != b[r - 1]);
            if (up + 1 < v) v = up + 1;
            if (D[i - 1] + 1 < v) v = D[i - 1] + 1;
            diag = up;
            D[i] = v;
        }
    }
}

/* levenshtein_align_hirschberg (Levenshtein_impl.hpp:1185-1212): strip the common affixes;
 * when the band's matrix stays under 1 MiB (or |s1| < 65 or |s2| < 10) walk the pair's
 * matrix, else split s2 in the middle at the s1 position find_hirschberg_pos
 * (Levenshtein_impl.hpp:1098-1182) picks -- the first position of the smallest sum of the
 * forward and backward distances (exact rows here; the reference's banded rows hold
 * every position that can reach that minimum) -- and recurse on both halves.
 * Returns the operations written at ops + opos. */
static size_t lev_hirschberg(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, size_t max, size_t src,
                             size_t dst, int32_t* type, uint64_t* src_pos, uint64_t* dest_pos, size_t opos)
{
    size_t pre = 0;
    while (pre < l1 && pre < l2 && s1[pre] == s2[pre]) ++pre;
    size_t suf = 0;
    while (suf < l1 - pre && suf < l2 - pre && s1[l1 - 1 - suf] == s2[l2 - 1 - suf]) ++suf;
    const uint32_t* a = s1 + pre;
    const uint32_t* b = s2 + pre;
    const size_t la = l1 - pre - suf, lb = l2 - pre - suf;
    src += pre;
    dst += pre;
    const size_t mx = la > lb ? la : lb;
    if (mx < max) max = mx;
    const size_t full_band = la < 2 * max + 1 ? la : 2 * max + 1;
    if ((2 * full_band * lb) / 8 < 1024 * 1024 || la < 65 || lb < 10)
        return ed_table_walk(ORC_LEVENSHTEIN, a, la, b, lb, src, dst, type + opos, src_pos + opos, dest_pos + opos);
    const size_t ls = lb / 2;
    uint64_t* L = (uint64_t*)malloc((la + 1) * sizeof(uint64_t));
    uint64_t* R = (uint64_t*)malloc((la + 1) * sizeof(uint64_t));
    uint32_t* ra = (uint32_t*)malloc((la + 1) * sizeof(uint32_t));
    uint32_t* rb = (uint32_t*)malloc((lb - ls + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < la; ++i) ra[i] = a[la - 1 - i];
    for (size_t i = 0; i < lb - ls; ++i) rb[i] = b[lb - 1 - i];
    lev_last_row(a, la, b, ls, L);
    lev_last_row(ra, la, rb, lb - ls, R); /* R[j] = distance(a[la - j, la), b[ls, lb)) */
    size_t mid = 0;
    uint64_t best = ~0ull;
    for (size_t i = 0; i <= la; ++i) {
        const uint64_t t = L[i] + R[la - i];
        if (t < best) best = t, mid = i;
    }
    const size_t left = (size_t)L[mid], right = (size_t)R[la - mid];
    free(L);
    free(R);
    free(ra);
    free(rb);
    const size_t k1 = lev_hirschberg(a, mid, b, ls, left, src, dst, type, src_pos, dest_pos, opos);
    const size_t k2 =
        lev_hirschberg(a + mid, la - mid, b + ls, lb - ls, right, src + mid, dst + ls, type, src_pos, dest_pos, opos + k1);
    return k1 + k2;
}

size_t orc_editops(int metric, const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, int32_t* type,
                   uint64_t* src_pos, uint64_t* dest_pos)
{
    if (metric == ORC_HAMMING) {
        /* hamming_editops with pad (Hamming_impl.hpp:40-59): a replacement at every
         * differing position of the common length, then the longer string's tail */
        size_t k = 0, i = 0;
        const size_t mn = l1 < l2 ? l1 : l2;
        for (; i < mn; ++i)
            if (s1[i] != s2[i]) type[k] = ED_REPLACE, src_pos[k] = i, dest_pos[k] = i, ++k;
        for (; i < l1; ++i) type[k] = ED_DELETE, src_pos[k] = i, dest_pos[k] = l2, ++k;
        for (; i < l2; ++i) type[k] = ED_INSERT, src_pos[k] = l1, dest_pos[k] = i, ++k;
        return k;
    }
    size_t pre = 0;
    while (pre < l1 && pre < l2 && s1[pre] == s2[pre]) ++pre;
    size_t suf = 0;
    while (suf < l1 - pre && suf < l2 - pre && s1[l1 - 1 - suf] == s2[l2 - 1 - suf]) ++suf;
    const uint32_t* a = s1 + pre;
    const uint32_t* b = s2 + pre;
    const size_t la = l1 - pre - suf, lb = l2 - pre - suf;
    if (metric == ORC_LEVENSHTEIN && la > 64 && lb >= 10 && (2 * la * lb) / 8 >= 1024 * 1024)
        return lev_hirschberg(s1, l1, s2, l2, l1 > l2 ? l1 : l2, 0, 0, type, src_pos, dest_pos, 0);
    if (metric == ORC_LEVENSHTEIN && la > 64 && lb > 0 && g_lev_block)
        return lev_block_editops(a, la, b, lb, pre, type, src_pos, dest_pos);
    return ed_table_walk(metric, a, la, b, lb, pre, pre, type, src_pos, dest_pos);
}

/* damerau_levenshtein_distance (DamerauLevenshtein_impl.hpp:105-121): the length
 * difference bound, then the exact distance, cut. */
static uint64_t damerau_cached_distance(const Pair* p, uint64_t c)
{
    uint64_t min_edits = p->l1 > p->l2 ? p->l1 - p->l2 : p->l2 - p->l1;
    if (min_edits > c) return c + 1;
    uint64_t d = orc_damerau_raw(p->s1, p->l1, p->s2, p->l2);
    return d <= c ? d : c + 1;
}

/* CachedOSA::_distance (OSA.hpp:253-267). */
static uint64_t osa_cached_distance(const Pair* p, uint64_t c)
{
    uint64_t res;
    if (p->l1 == 0)
        res = p->l2;
    else if (p->l2 == 0)
        res = p->l1;
    else {
        res = orc_osa_raw(p->s1, p->l1, p->s2, p->l2);
        res = res <= c ? res : c + 1;
    }
    return (res <= c) ? res : c + 1;
}

static uint64_t int_maximum(const Pair* p)
{
    switch (p->metric) {
    case ORC_LEVENSHTEIN: return lev_maximum(p->l1, p->l2, p->ins, p->del, p->rep);
    case ORC_INDEL: return (uint64_t)p->l1 + (uint64_t)p->l2;
    case ORC_LCSSEQ:
    case ORC_OSA:
    case ORC_HAMMING:
    case ORC_PREFIX:
    case ORC_POSTFIX:
    case ORC_DAMERAU_LEVENSHTEIN: return u64max(p->l1, p->l2); /* DamerauLevenshtein.hpp:127-131 */
    }
    return 0;
}

/* similarity-type integer metrics (CachedSimilarityBase) */
static int is_sim_metric(int metric) { return metric == ORC_LCSSEQ || metric == ORC_PREFIX || metric == ORC_POSTFIX; }

static uint64_t sim_cached_similarity(const Pair* p, uint64_t c)
{
    return p->metric == ORC_LCSSEQ ? lcs_cached_similarity(p, c) : affix_cached_similarity(p, c);
}

/* derived._distance for distance-type integer metrics */
static uint64_t int_distance_core(const Pair* p, uint64_t c)
{
    switch (p->metric) {
    case ORC_LEVENSHTEIN: return lev_cached_distance(p, c);
    case ORC_INDEL: return indel_cached_distance(p, c);
    case ORC_OSA: return osa_cached_distance(p, c);
    case ORC_HAMMING: return hamming_cached_distance(p, c);
    case ORC_DAMERAU_LEVENSHTEIN: return damerau_cached_distance(p, c);
    }
    return 0;
}

/* CachedSimilarityBase::_distance for LCSseq, Prefix, Postfix (details/distance.hpp:359-369). */
static uint64_t lcs_cached_distance(const Pair* p, uint64_t c)
{
    uint64_t maximum = int_maximum(p);
    uint64_t cs = (maximum > c) ? maximum - c : 0;
    uint64_t sim = sim_cached_similarity(p, cs);
    uint64_t dist = maximum - sim;
    return (dist <= c) ? dist : c + 1;
}

static uint64_t int_api_distance(const Pair* p, uint64_t c)
{
    if (is_sim_metric(p->metric)) return lcs_cached_distance(p, c);
    return int_distance_core(p, c);
}

static uint64_t int_api_similarity(const Pair* p, uint64_t c)
{
    if (is_sim_metric(p->metric)) return sim_cached_similarity(p, c);
    /* CachedDistanceBase::_similarity (details/distance.hpp:306-317) */
    uint64_t maximum = int_maximum(p);
    if (c > maximum) return 0;
    uint64_t cd = maximum - c;
    uint64_t dist = int_distance_core(p, cd);
    uint64_t sim = maximum - dist;
    return (sim >= c) ? sim : 0;
}

/* CachedNormalizedMetricBase::_normalized_distance (details/distance.hpp:240-252). */
static double int_api_norm_distance(const Pair* p, double c)
{
    uint64_t maximum = int_maximum(p);
    uint64_t cd = (uint64_t)ceil((double)maximum * c);
    double dist = (double)int_api_distance(p, cd);
    double norm = (maximum != 0) ? dist / (double)maximum : 0.0;
    return (norm <= c) ? norm : 1.0;
}

/* details/common.hpp:50-53 */
static double normsim_to_normdist(double c) { return dmin(1.0, 1.0 - c + 0.00001); }

/* CachedNormalizedMetricBase::_normalized_similarity (details/distance.hpp:254-262). */
static double int_api_norm_similarity(const Pair* p, double c)
{
    double cs = normsim_to_normdist(c);
    double nd = int_api_norm_distance(p, cs);
    double ns = 1.0 - nd;
    return (ns >= c) ? ns : 0.0;
}

/* Jaro / JaroWinkler: CachedSimilarityBase<double> with maximum 1.0. */
static double fp_similarity(const Pair* p, double c)
{
    if (p->metric == ORC_JARO) return orc_jaro_similarity(p->s1, p->l1, p->s2, p->l2, c);
    return orc_jaro_winkler_similarity(p->s1, p->l1, p->s2, p->l2, p->pw, c);
}

static double fp_distance(const Pair* p, double c)
{
    double maximum = 1.0;
    double cs = (maximum > c) ? maximum - c : 0;
    double sim = fp_similarity(p, cs);
    double dist = maximum - sim;
    return (dist <= c) ? dist : 1.0;
}

static double fp_norm_distance(const Pair* p, double c)
{
    double maximum = 1.0;
    double cd = ceil(maximum * c);
    double dist = fp_distance(p, cd);
    double norm = (maximum != 0) ? dist / maximum : 0.0;
    return (norm <= c) ? norm : 1.0;
}

static double fp_norm_similarity(const Pair* p, double c)
{
    double cs = normsim_to_normdist(c);
    double nd = fp_norm_distance(p, cs);
    double ns = 1.0 - nd;
    return (ns >= c) ? ns : 0.0;
}

/* ------------------------------------------------------------------------ */
/* fuzz::partial_ratio (fuzz_impl.hpp:57-279)                                */
/* ------------------------------------------------------------------------ */

/* CachedRatio(s1).similarity(s2, cutoff) (fuzz_impl.hpp:42-47) */
static double ratio_similarity(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff)
{
    Pair p = {ORC_INDEL, s1, s2, l1, l2, 1, 1, 2, 0.0};
    return int_api_norm_similarity(&p, cutoff / 100) * 100;
}

/* CachedIndel(s1).distance(s2) with the default (unbounded) cutoff */
static int cmp_u32(const void* x, const void* y)
{
    const uint32_t a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    return a < b ? -1 : (a > b);
}

/* The LCS length by the bit-parallel recurrence (Hyyro 2004; the reference's
 * LCSseq_impl.hpp:205-250): S starts all ones, each unit of s2 with match words X
 * (bit i set where s1[i] equals it) does U = S & X, S = (S + U) | (S - U) over the
 * words with the addition's carries, and the LCS is the number of clear bits.  The
 * partial ratios' many windows use it; tests/test_oracle_vs_reference.py checks it
 * against the DP of orc_lcs_raw. */
uint64_t orc_lcs_bits(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    if (!l1 || !l2) return 0;
    const size_t W = (l1 + 63) / 64;
    uint32_t* sym = (uint32_t*)malloc(l1 * sizeof(uint32_t));
    memcpy(sym, s1, l1 * sizeof(uint32_t));
    qsort(sym, l1, sizeof(uint32_t), cmp_u32);
    size_t d = 0;
    for (size_t i = 0; i < l1; ++i)
        if (i == 0 || sym[i] != sym[d - 1]) sym[d++] = sym[i];
    uint64_t* pm = (uint64_t*)calloc(d * W, sizeof(uint64_t));
    for (size_t i = 0; i < l1; ++i) {
        const uint32_t* k = (const uint32_t*)bsearch(&s1[i], sym, d, sizeof(uint32_t), cmp_u32);
        pm[(size_t)(k - sym) * W + i / 64] |= 1ull << (i % 64);
    }
    uint64_t* S = (uint64_t*)malloc(W * sizeof(uint64_t));
    for (size_t w = 0; w < W; ++w) S[w] = ~0ull;
    for (size_t j = 0; j < l2; ++j) {
        const uint32_t* k = (const uint32_t*)bsearch(&s2[j], sym, d, sizeof(uint32_t), cmp_u32);
        if (!k) continue; /* no match: S is unchanged */
        const uint64_t* X = pm + (size_t)(k - sym) * W;
        uint64_t carry = 0;
        for (size_t w = 0; w < W; ++w) {
            const uint64_t u = S[w] & X[w];
            const uint64_t x = S[w] + u + carry;
            carry = (x < S[w] || (carry && x == S[w])) ? 1 : 0;
            S[w] = x | (S[w] - u);
        }
    }
    uint64_t lcs = 0;
    for (size_t w = 0; w < W; ++w) {
        const size_t bits = l1 - 64 * w >= 64 ? 64 : l1 - 64 * w;
        const uint64_t mask = bits == 64 ? ~0ull : ((1ull << bits) - 1);
        lcs += (uint64_t)__builtin_popcountll(~S[w] & mask);
    }
    free(sym);
    free(pm);
    free(S);
    return lcs;
}

static size_t indel_distance_full(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2)
{
    return l1 + l2 - 2 * (size_t)orc_lcs_bits(s1, l1, s2, l2);
}

/* CharSet::find of s1's units */
static int unit_in(const uint32_t* s, size_t l, uint32_t c)
{
    for (size_t i = 0; i < l; ++i)
        if (s[i] == c) return 1;
    return 0;
}

typedef struct {
    size_t first, second;
} Window;

/* fuzz_detail::partial_ratio_impl (fuzz_impl.hpp:68-179): s1 is the needle, len1 <= len2.
 * The full windows s2[p, p + len1), p < len2 - len1, are searched by bisection of
 * [0, len2 - len1 - 1] with the reference's bound on the distances between two
 * evaluated windows; then the prefix and suffix windows.  *ds, *de: the window of s2
 * the result reports (res.dest_start / dest_end: [0, len1) until a window improves;
 * a full window that beats the cutoff distance moves it even when its score then
 * falls below score_cutoff). */
static double partial_ratio_impl_at(const uint32_t* s1, size_t len1, const uint32_t* s2, size_t len2,
                                    double score_cutoff, size_t* ds, size_t* de)
{
    double res = 0;
    *ds = 0;
    *de = len1;
    if (len2 > len1) {
        const size_t maximum = len1 * 2;
        const double norm_cutoff_sim = normsim_to_normdist(score_cutoff / 100);
        size_t cutoff_dist = (size_t)ceil((double)maximum * norm_cutoff_sim);
        size_t best_dist = SIZE_MAX;
        const size_t nscores = len2 - len1;
        size_t* scores = (size_t*)malloc(nscores * sizeof(size_t));
        Window* windows = (Window*)malloc(2 * nscores * sizeof(Window) + sizeof(Window));
        Window* next = (Window*)malloc(2 * nscores * sizeof(Window) + sizeof(Window));
        size_t nwin = 1, nnext = 0;
        for (size_t i = 0; i < nscores; ++i) scores[i] = SIZE_MAX;
        windows[0].first = 0;
        windows[0].second = len2 - len1 - 1;
        int perfect = 0;
        while (nwin && !perfect) {
            for (size_t k = 0; k < nwin && !perfect; ++k) {
                const size_t f = windows[k].first, sc = windows[k].second;
                if (scores[f] == SIZE_MAX) {
                    scores[f] = indel_distance_full(s1, len1, s2 + f, len1);
                    if (scores[f] < cutoff_dist) {
                        cutoff_dist = best_dist = scores[f];
                        *ds = f;
                        *de = f + len1;
                        if (best_dist == 0) perfect = 1;
                    }
                }
                if (perfect) break;
                if (scores[sc] == SIZE_MAX) {
                    scores[sc] = indel_distance_full(s1, len1, s2 + sc, len1);
                    if (scores[sc] < cutoff_dist) {
                        cutoff_dist = best_dist = scores[sc];
                        *ds = sc;
                        *de = sc + len1;
                        if (best_dist == 0) perfect = 1;
                    }
                }
                if (perfect) break;
                const size_t cell_diff = sc - f;
                if (cell_diff == 1) continue;
                const size_t known_edits = scores[f] > scores[sc] ? scores[f] - scores[sc] : scores[sc] - scores[f];
                const size_t max_score_improvement = (cell_diff - known_edits / 2) / 2 * 2;
                const ptrdiff_t min_score =
                    (ptrdiff_t)(scores[f] < scores[sc] ? scores[f] : scores[sc]) - (ptrdiff_t)max_score_improvement;
                if (min_score < (ptrdiff_t)cutoff_dist) {
                    const size_t center = cell_diff / 2;
                    next[nnext].first = f;
                    next[nnext].second = f + center;
                    ++nnext;
                    next[nnext].first = f + center;
                    next[nnext].second = sc;
                    ++nnext;
                }
            }
            Window* t = windows;
            windows = next;
            next = t;
            nwin = nnext;
            nnext = 0;
        }
        free(scores);
        free(windows);
        free(next);
        if (perfect) return 100;
        double score = 1.0 - ((double)best_dist / (double)maximum);
        score *= 100;
        if (score >= score_cutoff) score_cutoff = res = score;
    }
    for (size_t i = 1; i < len1; ++i) {
        if (!unit_in(s1, len1, s2[i - 1])) continue;
        const double ls = ratio_similarity(s1, len1, s2, i, score_cutoff);
        if (ls > res) {
            score_cutoff = res = ls;
            *ds = 0;
            *de = i;
            if (res == 100.0) return res;
        }
    }
    for (size_t i = len2 - len1; i < len2; ++i) {
        if (!unit_in(s1, len1, s2[i])) continue;
        const double ls = ratio_similarity(s1, len1, s2 + i, len2 - i, score_cutoff);
        if (ls > res) {
            score_cutoff = res = ls;
            *ds = i;
            *de = len2;
            if (res == 100.0) return res;
        }
    }
    return res;
}

static double partial_ratio_impl(const uint32_t* s1, size_t len1, const uint32_t* s2, size_t len2,
                                 double score_cutoff)
{
    size_t ds, de;
    return partial_ratio_impl_at(s1, len1, s2, len2, score_cutoff, &ds, &de);
}

/* ------------------------------------------------------------------------ */
/* token sort (common_impl.hpp:84-179, SplittedSentenceView.hpp:70-85)      */
/* ------------------------------------------------------------------------ */

static int g_unit_width = 4;
static int g_signed8 = 0;

void orc_set_units(int width, int signed8)
{
    g_unit_width = width;
    g_signed8 = signed8;
}

/* detail::is_space for the configured unit width */
static int orc_is_space(uint32_t c)
{
    switch (c) {
    case 0x0009: case 0x000A: case 0x000B: case 0x000C: case 0x000D:
    case 0x001C: case 0x001D: case 0x001E: case 0x001F: case 0x0020: return 1;
    }
    if (g_unit_width == 1) return 0;
    switch (c) {
    case 0x0085: case 0x00A0: case 0x1680: case 0x2000: case 0x2001: case 0x2002: case 0x2003:
    case 0x2004: case 0x2005: case 0x2006: case 0x2007: case 0x2008: case 0x2009: case 0x200A:
    case 0x2028: case 0x2029: case 0x202F: case 0x205F: case 0x3000: return 1;
    }
    return 0;
}

static long long unit_key(uint32_t c) { return (g_unit_width == 1 && g_signed8) ? (long long)(signed char)c : (long long)c; }

/* Range operator< (Range.hpp:214-218): std::lexicographical_compare of the units */
static int tok_less(const uint32_t* a, size_t la, const uint32_t* b, size_t lb)
{
    size_t l = la < lb ? la : lb;
    for (size_t k = 0; k < l; ++k) {
        long long x = unit_key(a[k]), y = unit_key(b[k]);
        if (x != y) return x < y;
    }
    return la < lb;
}

size_t orc_sorted_join(const uint32_t* s, size_t n, uint32_t* out)
{
    size_t* st = (size_t*)malloc((n + 1) * sizeof(size_t));
    size_t* ln = (size_t*)malloc((n + 1) * sizeof(size_t));
    size_t T = 0;
    /* the split loop: tokens are the runs of non-space units */
    size_t i = 0;
    while (i < n) {
        while (i < n && orc_is_space(s[i])) ++i;
        if (i == n) break;
        size_t e = i;
        while (e < n && !orc_is_space(s[e])) ++e;
        st[T] = i;
        ln[T] = e - i;
        ++T;
        i = e;
    }
    /* std::sort: equal tokens are identical, so any correct sort gives the same join */
    for (size_t a = 1; a < T; ++a) {
        size_t ks = st[a], kl = ln[a], b = a;
        while (b > 0 && tok_less(s + ks, kl, s + st[b - 1], ln[b - 1])) {
            st[b] = st[b - 1];
            ln[b] = ln[b - 1];
            --b;
        }
        st[b] = ks;
        ln[b] = kl;
    }
    size_t len = 0;
    for (size_t t = 0; t < T; ++t) {
        if (t) out[len++] = 0x20;
        memcpy(out + len, s + st[t], ln[t] * sizeof(uint32_t));
        len += ln[t];
    }
    free(st);
    free(ln);
    return len;
}

/* ------------------------------------------------------------------------ */
/* token sets (common_impl.hpp:11-35, fuzz_impl.hpp:370-418, 518-565)       */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint32_t* p;
    size_t n;
} Tok;

/* sorted_split: the tokens in sorted order; returns the count */
static size_t sorted_tokens(const uint32_t* s, size_t n, Tok* out)
{
    size_t T = 0, i = 0;
    while (i < n) {
        while (i < n && orc_is_space(s[i])) ++i;
        if (i == n) break;
        size_t e = i;
        while (e < n && !orc_is_space(s[e])) ++e;
        out[T].p = s + i;
        out[T].n = e - i;
        ++T;
        i = e;
    }
    for (size_t a = 1; a < T; ++a) {
        Tok k = out[a];
        size_t b = a;
        while (b > 0 && tok_less(k.p, k.n, out[b - 1].p, out[b - 1].n)) {
            out[b] = out[b - 1];
            --b;
        }
        out[b] = k;
    }
    return T;
}

static int tok_eq(Tok a, Tok b) { return a.n == b.n && (a.n == 0 || memcmp(a.p, b.p, a.n * sizeof(uint32_t)) == 0); }

/* SplittedSentenceView::dedupe (std::unique) */
static size_t dedupe(Tok* t, size_t T)
{
    size_t w = 0;
    for (size_t i = 0; i < T; ++i)
        if (w == 0 || !tok_eq(t[w - 1], t[i])) t[w++] = t[i];
    return w;
}

/* SplittedSentenceView::size() of a token list: units plus one space between tokens */
static size_t joined_len(const Tok* t, size_t T)
{
    size_t n = T ? T - 1 : 0;
    for (size_t i = 0; i < T; ++i) n += t[i].n;
    return n;
}

static size_t join_into(const Tok* t, size_t T, uint32_t* out)
{
    size_t len = 0;
    for (size_t i = 0; i < T; ++i) {
        if (i) out[len++] = 0x20;
        memcpy(out + len, t[i].p, t[i].n * sizeof(uint32_t));
        len += t[i].n;
    }
    return len;
}

/* fuzz_detail::norm_distance (fuzz_impl.hpp:55-61) */
static double fz_norm_distance(size_t dist, size_t lensum, double score_cutoff)
{
    double score = (lensum > 0) ? (100.0 - 100.0 * (double)dist / (double)lensum) : 100.0;
    return (score >= score_cutoff) ? score : 0;
}

/* The set part shared by token_set_ratio and token_ratio.  *perfect: one token set
 * holds the other (and they share a token).  Returns max(result of the diff
 * joins, sect+ab / sect+ba ratios); ratio_in seeds `result` (token_ratio passes the
 * sorted-join ratio, token_set_ratio 0). */
static double token_set_part(Tok* ta, size_t na, Tok* tb, size_t nb, double score_cutoff, double ratio_in,
                             int* perfect)
{
    /* set_decomposition (common_impl.hpp:11-35) */
    na = dedupe(ta, na);
    nb = dedupe(tb, nb);
    Tok* inter = (Tok*)malloc((na + 1) * sizeof(Tok));
    Tok* dab = (Tok*)malloc((na + 1) * sizeof(Tok));
    Tok* dba = (Tok*)malloc((nb + 1) * sizeof(Tok));
    size_t ni = 0, nab = 0, nba = nb;
    memcpy(dba, tb, nb * sizeof(Tok));
    for (size_t i = 0; i < na; ++i) {
        size_t j = 0;
        while (j < nba && !tok_eq(dba[j], ta[i])) ++j;
        if (j < nba) {
            memmove(dba + j, dba + j + 1, (nba - j - 1) * sizeof(Tok));
            --nba;
            inter[ni++] = ta[i];
        }
        else {
            dab[nab++] = ta[i];
        }
    }
    *perfect = ni && (!nab || !nba);
    double result = ratio_in;
    if (!*perfect) {
        const size_t ab_len = joined_len(dab, nab), ba_len = joined_len(dba, nba), sect_len = joined_len(inter, ni);
        uint32_t* a = (uint32_t*)malloc((ab_len + 1) * sizeof(uint32_t));
        uint32_t* b = (uint32_t*)malloc((ba_len + 1) * sizeof(uint32_t));
        join_into(dab, nab, a);
        join_into(dba, nba, b);
        const size_t sect_ab_len = sect_len + (sect_len != 0) + ab_len;
        const size_t sect_ba_len = sect_len + (sect_len != 0) + ba_len;
        /* score_cutoff_to_distance (fuzz_impl.hpp:63-66) */
        const size_t cutoff_distance =
            (size_t)ceil((double)(sect_ab_len + sect_ba_len) * (1.0 - score_cutoff / 100));
        /* indel_distance(diff_ab_joined, diff_ba_joined, cutoff_distance) */
        size_t dist = ab_len + ba_len - 2 * (size_t)orc_lcs_raw(a, ab_len, b, ba_len);
        if (dist > cutoff_distance) dist = cutoff_distance + 1;
        if (dist <= cutoff_distance) {
            const double v = fz_norm_distance(dist, sect_ab_len + sect_ba_len, score_cutoff);
            result = result < v ? v : result;
        }
        if (sect_len) {
            const double r1 = fz_norm_distance((sect_len != 0) + ab_len, sect_len + sect_ab_len, score_cutoff);
            const double r2 = fz_norm_distance((sect_len != 0) + ba_len, sect_len + sect_ba_len, score_cutoff);
            result = result < r1 ? r1 : result; /* std::max({result, r1, r2}) */
            result = result < r2 ? r2 : result;
        }
        free(a);
        free(b);
    }
    free(inter);
    free(dab);
    free(dba);
    return result;
}

/* fuzz::token_set_ratio (fuzz_impl.hpp:370-426) */
static double orc_token_set_ratio(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff)
{
    if (cutoff > 100) return 0;
    Tok* ta = (Tok*)malloc((l1 + 1) * sizeof(Tok));
    Tok* tb = (Tok*)malloc((l2 + 1) * sizeof(Tok));
    size_t na = sorted_tokens(s1, l1, ta), nb = sorted_tokens(s2, l2, tb);
    double r = 0;
    if (na && nb) {
        int perfect = 0;
        r = token_set_part(ta, na, tb, nb, cutoff, 0.0, &perfect);
        if (perfect) r = 100;
    }
    free(ta);
    free(tb);
    return r;
}

/* fuzz::token_ratio / CachedTokenRatio::similarity (fuzz_impl.hpp:518-565, 574-622, 683-691) */
static double orc_token_ratio(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff)
{
    if (cutoff > 100) return 0;
    Tok* ta = (Tok*)malloc((l1 + 1) * sizeof(Tok));
    Tok* tb = (Tok*)malloc((l2 + 1) * sizeof(Tok));
    size_t na = sorted_tokens(s1, l1, ta), nb = sorted_tokens(s2, l2, tb);
    /* ratio(tokens_a.join(), tokens_b.join()) before the tokens are deduplicated */
    uint32_t* a = (uint32_t*)malloc((l1 + 1) * sizeof(uint32_t));
    uint32_t* b = (uint32_t*)malloc((l2 + 1) * sizeof(uint32_t));
    size_t la = join_into(ta, na, a), lb = join_into(tb, nb, b);
    Pair q = {ORC_INDEL, a, b, la, lb, 1, 1, 2, 0.0};
    const double ratio = int_api_norm_similarity(&q, cutoff / 100) * 100;
    int perfect = 0;
    double r = token_set_part(ta, na, tb, nb, cutoff, ratio, &perfect);
    if (perfect) r = 100;
    free(a);
    free(b);
    free(ta);
    free(tb);
    return r;
}

/* fuzz_detail::partial_token_set_ratio (fuzz_impl.hpp:460-478) behind CachedPartialTokenSetRatio
 * (fuzz_impl.hpp:497-505) */
static double orc_partial_token_set_ratio(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2,
                                          double cutoff)
{
    if (cutoff > 100) return 0;
    Tok* ta = (Tok*)malloc((l1 + 1) * sizeof(Tok));
    Tok* tb = (Tok*)malloc((l2 + 1) * sizeof(Tok));
    size_t na = sorted_tokens(s1, l1, ta), nb = sorted_tokens(s2, l2, tb);
    double r = 0;
    if (na && nb) {
        na = dedupe(ta, na);
        nb = dedupe(tb, nb);
        /* set_decomposition: any common token -> 100, else the differences are the sets */
        int common = 0;
        for (size_t i = 0; i < na && !common; ++i)
            for (size_t j = 0; j < nb && !common; ++j) common = tok_eq(ta[i], tb[j]);
        if (common) {
            r = 100;
        }
        else {
            uint32_t* a = (uint32_t*)malloc((l1 + 1) * sizeof(uint32_t));
            uint32_t* b = (uint32_t*)malloc((l2 + 1) * sizeof(uint32_t));
            size_t la = join_into(ta, na, a), lb = join_into(tb, nb, b);
            r = orc_partial_ratio(a, la, b, lb, cutoff);
            free(a);
            free(b);
        }
    }
    free(ta);
    free(tb);
    return r;
}

/* fuzz_detail::partial_token_ratio (fuzz_impl.hpp:740-770) behind CachedPartialTokenRatio */
static double orc_partial_token_ratio(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff)
{
    if (cutoff > 100) return 0;
    Tok* ta = (Tok*)malloc((l1 + 1) * sizeof(Tok));
    Tok* tb = (Tok*)malloc((l2 + 1) * sizeof(Tok));
    Tok* da = (Tok*)malloc((l1 + 1) * sizeof(Tok));
    Tok* db = (Tok*)malloc((l2 + 1) * sizeof(Tok));
    const size_t na = sorted_tokens(s1, l1, ta), nb = sorted_tokens(s2, l2, tb);
    memcpy(da, ta, na * sizeof(Tok));
    memcpy(db, tb, nb * sizeof(Tok));
    const size_t nda = dedupe(da, na), ndb = dedupe(db, nb);
    int common = 0;
    for (size_t i = 0; i < nda && !common; ++i)
        for (size_t j = 0; j < ndb && !common; ++j) common = tok_eq(da[i], db[j]);
    double r;
    if (common) {
        r = 100;
    }
    else {
        uint32_t* a = (uint32_t*)malloc((l1 + 1) * sizeof(uint32_t));
        uint32_t* b = (uint32_t*)malloc((l2 + 1) * sizeof(uint32_t));
        size_t la = join_into(ta, na, a), lb = join_into(tb, nb, b);
        r = orc_partial_ratio(a, la, b, lb, cutoff); /* partial_ratio(s1_sorted, tokens_b.join()) */
        /* with no common token the differences are the deduplicated sets; when neither side
         * had duplicates they equal the sorted tokens and the reference skips the second call */
        if (!(na == nda && nb == ndb)) {
            const double c2 = (cutoff < r) ? r : cutoff; /* std::max(score_cutoff, result) */
            la = join_into(da, nda, a);
            lb = join_into(db, ndb, b);
            const double r2 = orc_partial_ratio(a, la, b, lb, c2);
            r = (r < r2) ? r2 : r;
        }
        free(a);
        free(b);
    }
    free(ta);
    free(tb);
    free(da);
    free(db);
    return r;
}

/* partial_ratio_alignment(...).score (fuzz_impl.hpp:193-228); CachedPartialRatio::similarity
 * (fuzz_impl.hpp:255-279) returns the same value. */
double orc_partial_ratio(const uint32_t* s1, size_t len1, const uint32_t* s2, size_t len2, double score_cutoff)
{
    if (len1 > len2) return orc_partial_ratio(s2, len2, s1, len1, score_cutoff);
    if (score_cutoff > 100) return 0;
    if (!len1 || !len2) return (len1 == len2) ? 100.0 : 0.0;
    const double score = partial_ratio_impl(s1, len1, s2, len2, score_cutoff);
    if (score != 100 && len1 == len2) {
        score_cutoff = (score_cutoff < score) ? score : score_cutoff; /* std::max */
        const double score2 = partial_ratio_impl(s2, len2, s1, len1, score_cutoff);
        if (score2 > score) return score2;
    }
    return score;
}

/* fuzz::partial_ratio_alignment (fuzz_impl.hpp:196-230): out[0] score; src_start, src_end,
 * dest_start, dest_end in al[0..3] (s1 is the source). */
void orc_partial_ratio_alignment(const uint32_t* s1, size_t len1, const uint32_t* s2, size_t len2,
                                 double score_cutoff, double* score, size_t* al)
{
    if (len1 > len2) {
        size_t t[4];
        orc_partial_ratio_alignment(s2, len2, s1, len1, score_cutoff, score, t);
        al[0] = t[2];
        al[1] = t[3];
        al[2] = t[0];
        al[3] = t[1];
        return;
    }
    al[0] = 0;
    al[1] = len1;
    al[2] = 0;
    al[3] = len1;
    if (score_cutoff > 100) {
        *score = 0;
        return;
    }
    if (!len1 || !len2) {
        *score = (len1 == len2) ? 100.0 : 0.0;
        return;
    }
    size_t ds, de;
    const double sc = partial_ratio_impl_at(s1, len1, s2, len2, score_cutoff, &ds, &de);
    *score = sc;
    al[2] = ds;
    al[3] = de;
    if (sc != 100 && len1 == len2) {
        score_cutoff = (score_cutoff < sc) ? sc : score_cutoff; /* std::max */
        size_t ds2, de2;
        const double sc2 = partial_ratio_impl_at(s2, len2, s1, len1, score_cutoff, &ds2, &de2);
        if (sc2 > sc) {
            /* the second pass's source is s2: swap its fields back */
            *score = sc2;
            al[0] = ds2;
            al[1] = de2;
            al[2] = 0;
            al[3] = len2;
        }
    }
}

int orc_cached_score(int metric, int kind, const uint32_t* s1, size_t l1, const uint32_t* s2,
                     size_t l2, uint64_t ins, uint64_t del, uint64_t rep, double prefix_weight,
                     uint64_t cutoff_u, double cutoff_d, uint64_t* out_u, double* out_d)
{
    Pair p = {metric, s1, s2, l1, l2, ins, del, rep, prefix_weight};
    switch (metric) {
    case ORC_LEVENSHTEIN:
    case ORC_INDEL:
    case ORC_LCSSEQ:
    case ORC_OSA:
    case ORC_HAMMING:
    case ORC_PREFIX:
    case ORC_POSTFIX:
    case ORC_DAMERAU_LEVENSHTEIN:
        if (metric != ORC_LEVENSHTEIN) p.ins = p.del = 1, p.rep = (metric == ORC_INDEL) ? 2 : 1;
        switch (kind) {
        case ORC_DISTANCE: *out_u = int_api_distance(&p, cutoff_u); return 0;
        case ORC_SIMILARITY: *out_u = int_api_similarity(&p, cutoff_u); return 0;
        case ORC_NORM_DISTANCE: *out_d = int_api_norm_distance(&p, cutoff_d); return 1;
        case ORC_NORM_SIMILARITY: *out_d = int_api_norm_similarity(&p, cutoff_d); return 1;
        }
        return -1;
    case ORC_JARO:
    case ORC_JARO_WINKLER:
        switch (kind) {
        case ORC_DISTANCE: *out_d = fp_distance(&p, cutoff_d); return 1;
        case ORC_SIMILARITY: *out_d = fp_similarity(&p, cutoff_d); return 1;
        case ORC_NORM_DISTANCE: *out_d = fp_norm_distance(&p, cutoff_d); return 1;
        case ORC_NORM_SIMILARITY: *out_d = fp_norm_similarity(&p, cutoff_d); return 1;
        }
        return -1;
    case ORC_RATIO:
        /* fuzz::CachedRatio::similarity (fuzz_impl.hpp:42-47) */
        if (kind != ORC_SIMILARITY) return -1;
        p.metric = ORC_INDEL;
        p.ins = p.del = 1;
        p.rep = 2;
        *out_d = int_api_norm_similarity(&p, cutoff_d / 100) * 100;
        return 1;
    case ORC_PARTIAL_RATIO:
        if (kind != ORC_SIMILARITY) return -1;
        *out_d = orc_partial_ratio(s1, l1, s2, l2, cutoff_d);
        return 1;
    case ORC_TOKEN_SORT_RATIO:
    case ORC_PARTIAL_TOKEN_SORT_RATIO: {
        /* CachedTokenSortRatio / CachedPartialTokenSortRatio::similarity (fuzz_impl.hpp:314-322,
         * 353-362): ratio / partial_ratio of the sorted-token joins */
        if (kind != ORC_SIMILARITY) return -1;
        if (cutoff_d > 100) {
            *out_d = 0;
            return 1;
        }
        uint32_t* a = (uint32_t*)malloc((l1 + 1) * sizeof(uint32_t));
        uint32_t* b = (uint32_t*)malloc((l2 + 1) * sizeof(uint32_t));
        size_t la = orc_sorted_join(s1, l1, a), lb = orc_sorted_join(s2, l2, b);
        if (metric == ORC_TOKEN_SORT_RATIO) {
            Pair q = {ORC_INDEL, a, b, la, lb, 1, 1, 2, 0.0};
            *out_d = int_api_norm_similarity(&q, cutoff_d / 100) * 100;
        }
        else {
            *out_d = orc_partial_ratio(a, la, b, lb, cutoff_d);
        }
        free(a);
        free(b);
        return 1;
    }
    case ORC_TOKEN_SET_RATIO:
        if (kind != ORC_SIMILARITY) return -1;
        *out_d = orc_token_set_ratio(s1, l1, s2, l2, cutoff_d);
        return 1;
    case ORC_TOKEN_RATIO:
        if (kind != ORC_SIMILARITY) return -1;
        *out_d = orc_token_ratio(s1, l1, s2, l2, cutoff_d);
        return 1;
    case ORC_PARTIAL_TOKEN_SET_RATIO:
        if (kind != ORC_SIMILARITY) return -1;
        *out_d = orc_partial_token_set_ratio(s1, l1, s2, l2, cutoff_d);
        return 1;
    case ORC_PARTIAL_TOKEN_RATIO:
        if (kind != ORC_SIMILARITY) return -1;
        *out_d = orc_partial_token_ratio(s1, l1, s2, l2, cutoff_d);
        return 1;
    case ORC_WRATIO: {
        /* CachedWRatio::similarity (fuzz_impl.hpp:838-876): ratio, then token_ratio (length
         * ratio < 1.5) or partial_ratio and partial_token_ratio, each with the cutoff raised */
        if (kind != ORC_SIMILARITY) return -1;
        double c = cutoff_d;
        if (c > 100 || !l1 || !l2) {
            *out_d = 0;
            return 1;
        }
        const double UNBASE_SCALE = 0.95;
        const double len_ratio = (l1 > l2) ? (double)l1 / (double)l2 : (double)l2 / (double)l1;
        p.metric = ORC_INDEL;
        p.ins = p.del = 1;
        p.rep = 2;
        double end_ratio = int_api_norm_similarity(&p, c / 100) * 100;
        if (len_ratio < 1.5) {
            c = std_max(c, end_ratio) / UNBASE_SCALE;
            const double r = orc_token_ratio(s1, l1, s2, l2, c);
            *out_d = std_max(end_ratio, r * UNBASE_SCALE);
            return 1;
        }
        const double PARTIAL_SCALE = (len_ratio <= 8.0) ? 0.9 : 0.6;
        c = std_max(c, end_ratio) / PARTIAL_SCALE;
        end_ratio = std_max(end_ratio, orc_partial_ratio(s1, l1, s2, l2, c) * PARTIAL_SCALE);
        c = std_max(c, end_ratio) / UNBASE_SCALE;
        const double r = orc_partial_token_ratio(s1, l1, s2, l2, c);
        *out_d = std_max(end_ratio, r * UNBASE_SCALE * PARTIAL_SCALE);
        return 1;
    }
    case ORC_QRATIO:
        /* CachedQRatio::similarity (fuzz_impl.hpp:912-922) */
        if (kind != ORC_SIMILARITY) return -1;
        if (!l1 || !l2) {
            *out_d = 0;
            return 1;
        }
        p.metric = ORC_INDEL;
        p.ins = p.del = 1;
        p.rep = 2;
        *out_d = int_api_norm_similarity(&p, cutoff_d / 100) * 100;
        return 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* experimental::Multi* API layer (details/distance.hpp:390-582)            */
/* ------------------------------------------------------------------------ */

/* Full (uncut) raw distance of the Multi kernels: Levenshtein/OSA SIMD kernels
 * report s2.size() for an empty s1 (Levenshtein_impl.hpp:393-394). */
static uint64_t multi_raw_distance(const Pair* p)
{
    switch (p->metric) {
    case ORC_LEVENSHTEIN:
        if (p->l1 == 0) return p->l2;
        return orc_levenshtein_raw(p->s1, p->l1, p->s2, p->l2, 1, 1, 1);
    case ORC_OSA:
        if (p->l1 == 0) return p->l2;
        return orc_osa_raw(p->s1, p->l1, p->s2, p->l2);
    case ORC_INDEL:
        /* MultiIndel::_distance (Indel.hpp:118-130) */
        return (uint64_t)p->l1 + p->l2 - 2 * orc_lcs_raw(p->s1, p->l1, p->s2, p->l2);
    case ORC_LCSSEQ:
        /* MultiSimilarityBase::_distance with the full similarity */
        return u64max(p->l1, p->l2) - orc_lcs_raw(p->s1, p->l1, p->s2, p->l2);
    }
    return 0;
}

static double multi_fp_similarity(const Pair* p, double c)
{
    if (p->metric == ORC_JARO) return orc_jaro_similarity(p->s1, p->l1, p->s2, p->l2, c);
    /* MultiJaroWinkler::_similarity (JaroWinkler.hpp:122-145) */
    double s = orc_jaro_similarity(p->s1, p->l1, p->s2, p->l2, dmin(0.7, c));
    if (s > 0.7) {
        size_t prefix = common_prefix4(p->s1, p->l1, p->s2, p->l2);
        s += (double)prefix * p->pw * (1.0 - s);
        s = dmin(s, 1.0);
    }
    if (s < c) s = 0.0;
    return s;
}

int orc_multi_score(int metric, int kind, const uint32_t* s1, size_t l1, const uint32_t* s2,
                    size_t l2, uint64_t ins, uint64_t del, uint64_t rep, double prefix_weight,
                    uint64_t cutoff_u, double cutoff_d, uint64_t* out_u, double* out_d)
{
    Pair p = {metric, s1, s2, l1, l2, 1, 1, 1, prefix_weight};
    if (metric == ORC_LEVENSHTEIN && (ins != 1 || del != 1 || rep > 2)) return -1;
    if (metric == ORC_LEVENSHTEIN || metric == ORC_OSA || metric == ORC_INDEL ||
        metric == ORC_LCSSEQ) {
        uint64_t maximum = (metric == ORC_INDEL) ? (uint64_t)l1 + l2 : u64max(l1, l2);
        if (metric == ORC_LEVENSHTEIN) maximum = lev_maximum(l1, l2, ins, del, rep);
        uint64_t d = multi_raw_distance(&p);
        if (metric == ORC_LCSSEQ && kind == ORC_SIMILARITY) {
            uint64_t lcs = maximum - d;
            *out_u = lcs >= cutoff_u ? lcs : 0;
            return 0;
        }
        switch (kind) {
        case ORC_DISTANCE: *out_u = d <= cutoff_u ? d : cutoff_u + 1; return 0;
        case ORC_SIMILARITY: {
            uint64_t sim = maximum - d;
            *out_u = sim >= cutoff_u ? sim : 0;
            return 0;
        }
        case ORC_NORM_DISTANCE: {
            double norm = maximum ? (double)d / (double)maximum : 0.0;
            *out_d = norm <= cutoff_d ? norm : 1.0;
            return 1;
        }
        case ORC_NORM_SIMILARITY: {
            double norm = maximum ? (double)d / (double)maximum : 0.0;
            norm = norm <= 1.0 ? norm : 1.0;
            double ns = 1.0 - norm;
            *out_d = ns >= cutoff_d ? ns : 0.0;
            return 1;
        }
        }
        return -1;
    }
    if (metric == ORC_JARO || metric == ORC_JARO_WINKLER) {
        switch (kind) {
        case ORC_SIMILARITY: *out_d = multi_fp_similarity(&p, cutoff_d); return 1;
        case ORC_DISTANCE: {
            double dist = 1.0 - multi_fp_similarity(&p, 0.0);
            *out_d = dist <= cutoff_d ? dist : 1.0;
            return 1;
        }
        case ORC_NORM_DISTANCE: {
            double dist = 1.0 - multi_fp_similarity(&p, 0.0);
            dist = dist <= 1.0 ? dist : 1.0;
            *out_d = dist <= cutoff_d ? dist : 1.0;
            return 1;
        }
        case ORC_NORM_SIMILARITY: {
            double dist = 1.0 - multi_fp_similarity(&p, 0.0);
            dist = dist <= 1.0 ? dist : 1.0;
            double ns = 1.0 - dist;
            *out_d = ns >= cutoff_d ? ns : 0.0;
            return 1;
        }
        }
        return -1;
    }
    if (metric == ORC_RATIO) {
        /* MultiRatio::similarity (fuzz.hpp:82-89) over MultiIndel */
        if (kind != ORC_SIMILARITY) return -1;
        p.metric = ORC_INDEL;
        uint64_t maximum = (uint64_t)l1 + l2;
        uint64_t d = multi_raw_distance(&p);
        double norm = maximum ? (double)d / (double)maximum : 0.0;
        norm = norm <= 1.0 ? norm : 1.0;
        double ns = 1.0 - norm;
        double c = cutoff_d / 100.0;
        *out_d = (ns >= c ? ns : 0.0) * 100.0;
        return 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* batch helpers                                                            */
/* ------------------------------------------------------------------------ */

int orc_cdist(int metric, int kind, int multi, const uint32_t* qchars, const uint64_t* qoff,
              size_t nq, const uint32_t* cchars, const uint64_t* coff, size_t nc, uint64_t ins,
              uint64_t del, uint64_t rep, double prefix_weight, uint64_t cutoff_u, double cutoff_d,
              uint64_t* out_u, double* out_d)
{
    int r = -1;
    for (size_t q = 0; q < nq; ++q) {
        const uint32_t* s1 = qchars + qoff[q];
        size_t l1 = (size_t)(qoff[q + 1] - qoff[q]);
        for (size_t c = 0; c < nc; ++c) {
            const uint32_t* s2 = cchars + coff[c];
            size_t l2 = (size_t)(coff[c + 1] - coff[c]);
            uint64_t u = 0;
            double d = 0;
            if (multi)
                r = orc_multi_score(metric, kind, s1, l1, s2, l2, ins, del, rep, prefix_weight, cutoff_u,
                                    cutoff_d, &u, &d);
            else
                r = orc_cached_score(metric, kind, s1, l1, s2, l2, ins, del, rep, prefix_weight,
                                     cutoff_u, cutoff_d, &u, &d);
            if (r < 0) return r;
            if (r == 0)
                out_u[q * nc + c] = u;
            else
                out_d[q * nc + c] = d;
        }
    }
    return r;
}

int orc_pairs(const int32_t* metrics, int kind, const uint32_t* achars, const uint64_t* aoff,
              const uint32_t* bchars, const uint64_t* boff, size_t n, uint64_t cutoff_u,
              double cutoff_d, uint64_t* out_u, double* out_d)
{
    int r = -1;
    for (size_t i = 0; i < n; ++i) {
        uint64_t u = 0;
        double d = 0;
        r = orc_cached_score(metrics[i], kind, achars + aoff[i], (size_t)(aoff[i + 1] - aoff[i]),
                             bchars + boff[i], (size_t)(boff[i + 1] - boff[i]), 1, 1, 1, 0.1, cutoff_u,
                             cutoff_d, &u, &d);
        if (r < 0) return r;
        if (out_u) out_u[i] = u;
        if (out_d) out_d[i] = d;
    }
    return r;
}
