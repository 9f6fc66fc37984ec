// rfa_internal.hpp -- kernel argument blocks and launch entry points (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rfa_types.hpp"

namespace rfa {

// Query-tile kernel (rfa_tile.hip): tiles of 64 queries of <= 64 units.
struct TileArgs {
    const void* qchars;
    const uint64_t* qoff;
    const uint32_t* qorder; // query indices, grouped 64 per tile
    uint32_t nq;            // entries of qorder
    uint32_t tile_base;     // first tile of this launch
    const void* cchars;
    const uint64_t* coff;
    const uint32_t* corder; // choices to score (nullptr: 0 .. ncsel-1)
    uint32_t ncsel;
    uint32_t choices_per_block;
    uint64_t out_stride; // number of choices in the set (row length of out)
    uint32_t rows_cap;   // PM rows reserved in LDS (>= distinct symbols of any tile + 1)
    uint32_t hash_bits;  // log2 slots of the LDS symbol hash (16/32-bit units)
    ScoreSpec spec;
    void* out;
    const uint64_t* qcut; // per-query integer cutoff (nullptr: spec.cutoff_u)
    uint32_t settle;      // the cutoffs may decide results by length alone: test before each recurrence
    uint32_t norm_epi;    // the normalized kinds' fast epilogue covers the launch (norm_epi_setup)
    uint32_t narrow_int;  // integer results stay below 2^32 (held as 32 bits)
    // per-pair double cutoff pcut[q * out_stride + c] in place of spec.cutoff_d (fuzz::WRatio's
    // raised cutoffs); nullptr: spec.cutoff_d.  Launches that set it run the general epilogue.
    const double* pcut;
};

// rfa_pairs_noalloc's workspace: a static device array per GPU (rfa_noalloc.hip).
// 48 MiB holds the plan of 2^20 pairs (per-pair bins and items, 128 workgroup
// histograms over the plan's bins, the scan's temporary storage).
constexpr size_t NOALLOC_SCRATCH_BYTES = 48ull << 20;
constexpr uint64_t NOALLOC_MAX_PAIRS = 1ull << 20;
constexpr uint32_t NOALLOC_MAX_DEVICES = 64;
// Device address of `device`'s workspace (taken once per device with that device current).
hipError_t noalloc_scratch_address(int device, void** ptr);

size_t tile_lds_bytes(int char_width, int word_bits, uint32_t rows_cap, uint32_t hash_bits);
hipError_t launch_tile(int algo, int char_width, int word_bits, const TileArgs& a, uint32_t ntiles,
                       uint32_t nblocks_x, hipStream_t s);

// Block kernel (rfa_block.hip): one wave per pair, pattern words across lanes.
// Item i pairs string item_a[i] of set A with item_b[i] of set B; the pattern
// (bit-vector side) is A's string unless item_swap[i] is set.
struct BlockArgs {
    const void* achars;
    const uint64_t* aoff;
    const void* bchars;
    const uint64_t* boff;
    const uint32_t* item_a;
    const uint32_t* item_b;
    const uint8_t* item_swap;  // nullptr: never swap
    const uint64_t* out_index; // output element per item (explicit mode)
    uint32_t cross_nb;         // 0: explicit items; else item k = (item_a[k / cross_nb], item_b[k % cross_nb])
    uint64_t out_stride;       // cross mode: out[a * out_stride + b]
    uint32_t nitems;
    uint32_t reuse_pattern;    // consecutive items with the same pattern reuse its PM
    uint32_t* queue;           // work counter, zero at launch
    uint32_t batch;            // items taken per queue grab (>= 1)
    uint32_t qgram;            // banded launches: run the q-gram count filter (some 7k + 7 < m)
    uint64_t* pm_scratch;      // global arenas (when the PM does not fit in LDS)
    uint64_t pm_scratch_words; // uint64 words per workgroup arena
    uint64_t* tflag_scratch;   // Jaro text flags, tflag_words per workgroup
    uint64_t tflag_words;
    uint32_t hash_bits;  // unused (each pattern sizes its own symbol hash)
    uint32_t lds_bytes;  // dynamic LDS per workgroup; patterns that need more use the global arena
    ScoreSpec spec;
    void* out;
    // device-planned mode: item k is item_packed[k] = pair | swap << 31 for k < *nitems_dev,
    // the metric of a pair is pair_metric[pair] and its result goes to out[pair]
    const uint32_t* item_packed;
    const uint32_t* nitems_dev;
    const uint8_t* pair_metric; // metric per pair (RFA_LEVENSHTEIN .. RFA_OSA)
    const uint64_t* qcut; // cross mode: integer cutoff per string of set A (nullptr: spec.cutoff_u)
    const double* pcut;   // cross mode: double cutoff per item at its output index (nullptr: spec.cutoff_d)
};

// Largest item count of one block-kernel launch: the 32-bit queue head runs past the
// last item by at most one batch (<= 1024 items) per workgroup and must not wrap.
constexpr uint64_t BLOCK_MAX_ITEMS = 0xffffffffull - (1ull << 26);

// arena bytes of the largest pattern (symbol map + PM rows); rows: distinct symbols + 1
// of the largest 8-bit pattern when known, 0 for the alphabet bound
uint64_t block_arena_bytes(int char_width, uint32_t max_m, uint32_t rows = 0);
hipError_t launch_block(int algo, int char_width, const BlockArgs& a, uint32_t nblocks, uint32_t slots, hipStream_t s);

// Banded Levenshtein of long 8-bit queries x a choice set (rfa_band.hip): one
// workgroup per (query, chunk of choices), a diagonal window of band_words(k)
// 64-bit words per pair, floor(64 / band_words(k)) pairs per wave.
constexpr uint32_t BAND_THREADS = 256;
constexpr uint32_t BAND_MAX_WORDS = 64;
__host__ __device__ inline uint32_t band_words(uint32_t k) { return (k + 2 + 63) / 64; }
__host__ __device__ inline uint32_t band_pad(uint32_t k) { return (k + 1 + 63) / 64 + 1; }
// window words of a pair with WPL words per lane (the last lane's words all run)
__host__ __device__ inline uint32_t band_window_words(uint32_t k, uint32_t wpl)
{
    return (band_words(k) + wpl - 1) / wpl * wpl;
}
// uint64 words of one zero-padded PM row of a query of m units under cutoff k
__host__ __device__ inline uint32_t band_row_words(uint32_t m, uint32_t k, uint32_t wpl)
{
    return band_pad(k) + (m + 63) / 64 + (k + 64 * band_window_words(k, wpl) + 63) / 64 + 2;
}
// dynamic LDS of one workgroup: the byte -> row map and `rows` PM rows
inline uint64_t band_lds_bytes(uint32_t m, uint32_t k, uint32_t rows, uint32_t wpl)
{
    return 1024 + (uint64_t)rows * band_row_words(m, k, wpl) * 8;
}
// pairs one wave runs: floor(64 / lanes per pair)
inline uint32_t band_pairs_per_wave(uint32_t k, uint32_t wpl) { return 64 / ((band_words(k) + wpl - 1) / wpl); }
struct BandArgs {
    const uint8_t* qchars;
    const uint64_t* qoff;
    const uint32_t* queries; // query per slot; workgroup b runs slot b / chunks_per_query
    const uint8_t* cchars;
    const uint64_t* coff;
    uint32_t nc;             // choices
    uint32_t chunk;          // choices per workgroup
    uint32_t chunks_per_query;
    uint64_t out_stride;
    ScoreSpec spec;          // kind RFA_DISTANCE, Levenshtein; band_k the raw cutoff
    void* out;               // out[q * out_stride + c]
    const uint64_t* qcut;    // integer cutoff per query (nullptr: spec.cutoff_u)
    unsigned long long* counters; // nullptr, or [word-steps, pairs] accumulated by the launch
    const uint8_t* wpl;      // words per lane (1 or 2) of each query slot
};
// multi: some window has several words; wpl2: some query runs two words per lane;
// threads: workgroup size (64 .. BAND_THREADS, whole waves)
hipError_t launch_band(const BandArgs& a, uint32_t nblocks, uint32_t lds_bytes, bool multi, bool wpl2,
                       uint32_t threads, hipStream_t s);

// Lane kernel (rfa_lane.hip): one pair per lane, per-lane 128-slot maps per 64-unit block.
struct LaneArgs {
    const void* achars;
    const uint64_t* aoff;
    const void* bchars;
    const uint64_t* boff;
    const uint32_t* item_a;      // nullptr: item index
    const uint32_t* item_b;      // nullptr: item index
    const uint8_t* item_swap;    // nullptr: pattern is always A's string
    const uint8_t* item_metric;  // nullptr: spec.metric for every item
    const uint64_t* out_index;   // nullptr: out[item]
    uint64_t nitems;
    uint64_t* arena;             // per-lane maps
    uint64_t arena_words;        // uint64 words per lane
    uint32_t gen_base;           // generation tags of this launch start above gen_base
    ScoreSpec spec;
    void* out;
    // device-planned mode: items item_packed[*begin_dev .. nitems), pair | swap << 31,
    // metric pair_metric[pair], result out[pair]
    const uint32_t* item_packed;
    const uint32_t* begin_dev;
    const uint32_t* end_dev; // nullptr: nitems
    const uint8_t* pair_metric; // metric per pair (RFA_LEVENSHTEIN .. RFA_OSA)
};

// Segmented kernel (rfa_seg.hip): pairs with patterns of <= 4096 units under the
// Levenshtein or LCS recurrence.  Ranges r = 0..15 of the plan: r < 8 holds the
// pairs with texts of >= chain_steps units (the long sequential chains), r >= 8
// the rest; class k = r % 8 runs SEG = 64 >> k lanes per pair for k < 7, and
// k = 7 one lane per pair whose pattern has at most 32 units (32-bit words).
// chain_steps is also the length from which the longer string becomes the
// pattern (tuned per context by rfa_pairs_autotune; SEG_CHAIN_STEPS by default).
constexpr uint32_t SEG_CLASSES = 8;
constexpr uint32_t SEG_HALF_CLASS = 7;
constexpr uint32_t SEG_CHAIN_STEPS = 512;
struct SegArgs {
    const void* achars;
    const uint64_t* aoff;
    const void* bchars;
    const uint64_t* boff;
    const uint32_t* item_packed; // planned items, pair | swap << 31
    const uint8_t* pair_metric; // metric per pair (RFA_LEVENSHTEIN .. RFA_OSA)
    const uint32_t* counters;    // range r: items [counters[3 + r], counters[4 + r])
    ScoreSpec spec;
    void* out;                   // out[pair]
};

// symbol bits the segmented kernel compares for a launch whose largest code unit is max_symbol
uint32_t seg_symbol_bits(int char_width, uint64_t max_symbol);
// grid: the resident workgroups of the chosen instantiation on num_cus CUs, at most max_blocks
hipError_t launch_seg(int char_width, uint32_t nb, const SegArgs& a, uint32_t num_cus, uint32_t max_blocks,
                      hipStream_t s);

// Positional metrics (rfa_posit.hip): Hamming / Prefix / Postfix cdist, query tiles
// of 64 against runs of 8 consecutive choices; out[q * out_stride + c].
struct PositArgs {
    const void* qchars;
    const uint64_t* qoff;
    uint32_t nq;
    const void* cchars;
    const uint64_t* coff;
    uint32_t nc;
    ScoreSpec spec;
    void* out;
    uint64_t out_stride;
    const uint64_t* qcut; // integer cutoff per query (nullptr: spec.cutoff_u)
    uint32_t q_base;      // first query of the launch (set by launch_posit)
};

hipError_t launch_posit(int char_width, const PositArgs& a, uint32_t num_cus, hipStream_t s);

// partial_ratio (rfa_partial.hip).  Query tiles of at most 64 units run the tile
// kernel (TileArgs, the tile plan of the scorer; double results); longer queries
// run one wave per (query, choice) pair with the windows across the lanes.
hipError_t launch_ptile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                        hipStream_t s);

struct PartialLongArgs {
    const void* qchars;
    const uint64_t* qoff;
    const uint32_t* qlist; // the long queries
    uint32_t nql;
    const void* cchars;
    const uint64_t* coff;
    uint32_t nc;
    uint64_t out_stride;
    double* out;             // out[q * out_stride + c]
    double cutoff;
    const double* pcut;      // per-pair cutoff [q * out_stride + c] (nullptr: cutoff)
    uint64_t* pm_scratch;    // global pattern arenas (when the pattern does not fit in LDS)
    uint64_t pm_scratch_words;
    uint32_t lds_bytes;      // 0: patterns live in pm_scratch
};

// Token-sorted copy of a string set (rfa_tokens.hip): out_len, ntok, out_off and tok_off
// hold count + 1 entries (out_len[count] and ntok[count] zeroed by the caller);
// tok_start / tok_len and out hold at least as many entries as the set has units.
struct TokenArgs {
    const void* chars;
    const uint64_t* off;
    uint64_t count;
    uint64_t* out_len;
    uint64_t* ntok;
    uint64_t* out_off;
    uint64_t* tok_off;
    uint32_t* tok_start;
    uint32_t* tok_len;
    void* out;
    void* scan_temp;
    size_t scan_bytes;
    int signed8; // 8-bit units compare as signed (the reference's `char`)
    uint32_t num_cus;
    // deduplicated form (token sets): dup (one byte per token), set_off (count + 1) and the
    // kept tokens' start (in the output string), length and hash, in sorted order
    int dedup;
    uint8_t* dup;
    uint64_t* set_off;
    uint32_t* set_start;
    uint32_t* set_len;
    uint64_t* set_hash;
};
size_t tokens_scan_bytes(uint64_t count);
hipError_t launch_tokens(int char_width, const TokenArgs& a, hipStream_t s);

// token_set_ratio / token_ratio (rfa_tset.hip): query tiles of deduplicated token
// joins (TileArgs over them) against choices' joins; the token tables give each
// string's tokens in sorted order (start in its join, length, hash).  mode 0 writes
// token_set_ratio; mode 1 combines the token-sort ratio already in out into token_ratio;
// modes 2 / 3 finish partial_token_set_ratio / partial_token_ratio from partial_ratio values.
struct TsetArgs {
    TileArgs t;
    const uint64_t* q_set_off;
    const uint32_t* q_set_start;
    const uint32_t* q_set_len;
    const uint64_t* q_set_hash;
    const uint64_t* c_set_off;
    const uint32_t* c_set_start;
    const uint32_t* c_set_len;
    const uint64_t* c_set_hash;
    int mode; // 0 token_set_ratio, 1 token_ratio, 2 partial_token_set_ratio, 3 partial_token_ratio
    uint32_t tok_lds_u64; // set by launch_tset
    const double* aux;    // mode 3: partial_ratio of the deduplicated joins (out holds the sorted joins')
};
hipError_t launch_scatter_cols(double* dst, uint64_t stride, const double* src, const uint32_t* cols, uint64_t ncols,
                               uint64_t nq, hipStream_t s);
hipError_t launch_scatter_rows(double* dst, uint64_t stride, const double* src, const uint32_t* rows, uint64_t nrows,
                               hipStream_t s);
hipError_t launch_tset(int char_width, int word_bits, TsetArgs a, uint32_t ntiles, uint32_t nblocks_x,
                       hipStream_t s);
// The token-set ratios for queries the tile plan does not take (deduplicated joins of
// 65..TSET_LONG_MAX units, or tiles of too many symbols): one wave per (query, choice),
// the query's join as a pattern of up to 64 words across the lanes (k_tset_long).
constexpr uint32_t TSET_LONG_MAX = 4096;
struct TsetLongArgs {
    PartialLongArgs p; // the long queries, the choices, out, cutoffs and the pattern arenas
    const uint64_t* q_set_off;
    const uint32_t* q_set_start;
    const uint32_t* q_set_len;
    const uint64_t* q_set_hash;
    const uint64_t* c_set_off;
    const uint32_t* c_set_start;
    const uint32_t* c_set_len;
    const uint64_t* c_set_hash;
    int mode;
    const double* aux;
};
hipError_t launch_tset_long(int char_width, const TsetLongArgs& a, uint32_t nblocks, hipStream_t s);

// fuzz::WRatio's combining steps over an nq x nc matrix (rfa_wratio.hip): stage 0 turns the
// ratio R into the first raised cutoffs C; stage 1 folds partial_ratio X into R and C for
// pairs with a length ratio >= 1.5; stage 2 combines into out (token_ratio on entry).
struct WRatioArgs {
    const uint64_t* qoff; // the raw queries' offsets
    const uint64_t* coff; // the raw choices' offsets
    uint32_t nq, nc;
    double cutoff;
    double* R;
    double* C;
    const double* X;
    double* out;
};
hipError_t launch_wratio(int stage, const WRatioArgs& a, uint32_t num_cus, hipStream_t s);

// Levenshtein with general weights (rfa_wlev.hip).  Query tiles of at most 64 units run
// the tile kernel (TileArgs); longer queries run one wave per (query, choice) pair with
// the pattern's rows across the lanes, up to WLEV_LONG_MAX rows (the choice becomes the
// pattern when the query is longer).  Every value of the recurrence must stay below 2^30.
hipError_t launch_wtile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                        hipStream_t s);

struct WlevLongArgs {
    const void* qchars;
    const uint64_t* qoff;
    const uint32_t* qlist; // the long queries
    uint32_t nql;
    const void* cchars;
    const uint64_t* coff;
    uint32_t nc;
    uint64_t out_stride;
    ScoreSpec spec;
    void* out;            // out[q * out_stride + c]
    const uint64_t* qcut; // integer cutoff per query (nullptr: spec.cutoff_u)
    int32_t* bnd;         // pairs of two strings over WLEV_LONG_MAX units: a boundary row of
    uint64_t bnd_stride;  // bnd_stride ints (> the longest choice) per wave of the launch
};
constexpr uint32_t WLEV_LONG_MAX = 1024; // pattern rows a wave holds (longer patterns run in strips)
hipError_t launch_wlong(int char_width, uint32_t cells_per_lane, const WlevLongArgs& a, uint32_t nblocks,
                        hipStream_t s);

// Damerau-Levenshtein (rfa_dl.hip): query tiles of at most 64 units (TileArgs, the
// scorer's tile plan), integer or double results like k_tile.
hipError_t launch_dltile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                         hipStream_t s);
// Longer queries: one wave per (query, 64 choices), one lane per pair, the arrays in
// scratch (3 * (max_rows + 1) * 64 words per block; nblocks slices).
struct DlLongArgs {
    const void* qchars;
    const uint64_t* qoff;
    const uint32_t* qlist;
    uint32_t nql;
    uint32_t max_rows; // longest query of qlist
    const void* cchars;
    const uint64_t* coff;
    uint64_t nc;
    uint64_t out_stride;
    ScoreSpec spec;
    void* out;
    const uint64_t* qcut;
    uint32_t* scratch;
};
hipError_t launch_dllong(int char_width, const DlLongArgs& a, uint32_t nblocks, hipStream_t s);

// Editops (rfa_editops.hip).  Pair i is (a[aoff[i], aoff[i+1]), b[boff[i], boff[i+1]));
// its rows start at rows[2 * (boff[i] - b_base)] (Levenshtein) or rows[row_base[i]] (LCS).  k_ed_rows (walk = false)
// fills info; k_ed_walk (walk = true) writes pair i's operations at ops[offsets[i]].
struct EdInfo {
    uint64_t pre;   // common prefix length
    uint32_t la;    // stripped lengths
    uint32_t lb;
    uint32_t count; // operations
    uint32_t flag;  // 1: past the supported lengths; 2: a Levenshtein pair for the block kernels
};
struct EdOp { // the layout of rfa_editop
    int32_t type;
    int32_t reserved;
    uint64_t src_pos;
    uint64_t dest_pos;
};
// A piece of a Levenshtein pair split by Hirschberg: s1[a0, a0 + la) against s2[b0, b0 + lb)
// of pair `pair`, its `dist` operations at ops[opos], positions offset by src / dst.
struct EdLeaf {
    uint32_t pair, la, lb, dist;
    uint64_t a0, b0, src, dst, opos;
};
// A split of such a piece (find_hirschberg_pos): in a0 / la / b0 / lb, out mid / left / right
struct EdSplit {
    uint32_t pair, la, lb, mid;
    uint64_t a0, b0;
    uint32_t left, right;
};
struct EditopsArgs {
    const void* achars;
    const void* bchars;
    const uint64_t* aoff;
    const uint64_t* boff;
    uint64_t b_base;
    uint64_t* rows;
    EdInfo* info;
    const uint64_t* offsets;
    EdOp* ops;
    uint64_t n;
    const uint64_t* row_base; // LCS: first row word of each pair (rows of ceil(la / 64) words)
    uint64_t* counts;         // n + 1 operation counts (the last one zero), scanned into offsets
    uint32_t* err;            // set when a pair exceeds the length limit
    uint64_t* need;           // Levenshtein block pairs: the largest row area (words) of one pair
    uint64_t* long_rows;      // per wave of k_ed_levl: long_stride words of rows
    uint64_t long_stride;
    uint64_t* hcount;         // Levenshtein pairs the reference splits by Hirschberg: their number
    uint32_t* hlist;          // and indices
    const EdLeaf* leaves;     // k_ed_levl's leaf mode (nullptr: the pairs of flag 2)
    uint64_t nleaves;
};
// Levenshtein editops of a stripped first string of 65..EDITOPS_LEV_BLOCK_MAX units run the
// reference's block recurrence with its Ukkonen band (Levenshtein_impl.hpp:641-850) when the
// reference records the whole band (no Hirschberg split, Levenshtein_impl.hpp:1185-1200)
constexpr uint32_t EDITOPS_LEV_BLOCK_MAX = 4096;
hipError_t launch_editops_lev_block(int char_width, const EditopsArgs& a, bool walk, uint32_t nblocks,
                                    hipStream_t s);
// find_hirschberg_pos of every split (one wave each): tmp holds 2 * (PBIG-sized) rows per wave
hipError_t launch_editops_split(int char_width, const EditopsArgs& a, EdSplit* splits, uint64_t nsplits,
                                uint32_t* tmp, uint64_t tmp_stride, uint32_t nblocks, hipStream_t s);
size_t editops_scan_bytes(uint64_t n);
// offsets[0..n] = exclusive prefix sum of counts[0..n] (counts[n] = 0)
hipError_t editops_offsets(const uint64_t* counts, uint64_t* offsets, uint64_t n, void* temp, size_t bytes,
                           hipStream_t s);
constexpr uint32_t EDITOPS_LCS_MAX = 4096; // stripped first string, LCS editops
hipError_t launch_editops(int char_width, const EditopsArgs& a, bool lev, bool walk, uint32_t nblocks,
                          hipStream_t s);
hipError_t launch_editops_hamming(int char_width, const EditopsArgs& a, bool walk, uint32_t nblocks, hipStream_t s);

// fuzz::partial_ratio_alignment of pairs (rfa_pralign.hip): one wave per pair whose
// shorter string (the needle, l1 units) has 1..PRALIGN_MAX units.  Each search works in
// a workspace (pralign_ws_bytes): the needle's match masks over the haystack (l2 units)
// and over the reversed last l1 units and the LCS of every candidate window, with the
// 8-bit symbol tables of short needles in the space the LCS values take later.  A workspace lives in LDS when it fits the launch's
// per-wave LDS size, else in the item's slice of scratch.
constexpr uint32_t PRALIGN_MAX = 1024;
constexpr uint32_t PRALIGN_WAVES = 1;       // waves (pairs) per workgroup: LDS is taken per wave
constexpr uint32_t PRALIGN_LDS_CAP = 16384; // per-wave LDS, bytes
struct PrAlignArgs {
    const void* achars;
    const uint64_t* aoff;
    const void* bchars;
    const uint64_t* boff;
    const uint32_t* items;       // pair of each item
    const uint64_t* scratch_off; // byte offset of each item's scratch slice (items outside LDS)
    uint8_t* scratch;
    uint32_t nitems;
    uint32_t lds_bytes;          // per-wave LDS of the launch (multiple of 8, at most PRALIGN_LDS_CAP)
    double cutoff;
    uint64_t* out;               // per pair, rfa_score_alignment: score bits, src_start, src_end,
                                 // dest_start, dest_end
};
// 8-bit needles of at most two words build their masks from per-symbol tables (in LDS only)
__host__ __device__ inline bool pralign_symbol_tables(int char_width, uint64_t l1)
{
    return char_width == 1 && l1 <= 128;
}
__host__ __device__ inline uint64_t pralign_ws_bytes(uint64_t l1, uint64_t l2, bool tables)
{
    const uint64_t w = (l1 + 63) / 64, f = l2 - l1;
    const uint64_t masks = 8 * w * (l2 + l1);
    const uint64_t rest = (4 * (f + 2 * l1) + 7) / 8 * 8;
    const uint64_t sym = tables ? 256 * 8 * w : 0;
    return masks + (rest > sym ? rest : sym);
}
// words: 1, 2, 4, 8 or 16 (the needle's 64-unit words, rounded up)
hipError_t launch_pralign(int char_width, uint32_t words, const PrAlignArgs& a, hipStream_t s);

// Needles of PRALIGN_MAX + 1 .. PBIG_MAX units (k_pbig, one wave per pair): the needle and
// its reversal as multi-word patterns (one word per lane), a row index per haystack unit,
// the reference's bisection replayed window by window, prefixes and suffixes recorded in
// one pass each.  Per-wave workspace in scratch (pbig_ws_bytes).
constexpr uint32_t PBIG_MAX = 4096;
__host__ __device__ inline uint64_t pbig_arena_bytes(int char_width, uint64_t m)
{
    const uint64_t alpha = char_width == 1 ? 256 : (char_width == 2 ? 65536 : 0xffffffffull);
    const uint64_t d = m < alpha ? m : alpha;
    uint64_t map = 1024;
    if (char_width != 1) {
        uint32_t b = 6;
        while ((1ull << b) < 2 * d) ++b;
        map = (1ull << b) * 12 + 8;
    }
    return (map + (d + 1) * ((m + 63) / 64) * 8 + 16 + 255) / 256 * 256;
}
// offsets (bytes) of the workspace parts: arena, reversed arena, reversed needle, row index,
// reversed row index, window distances, two range lists, prefix and suffix LCS; [9] = total
__host__ __device__ inline void pbig_layout(int char_width, uint64_t l1, uint64_t l2, uint64_t* o)
{
    const uint64_t F = l2 - l1, ab = pbig_arena_bytes(char_width, l1);
    auto r = [](uint64_t x) { return (x + 255) / 256 * 256; };
    o[0] = 0;
    o[1] = ab;
    o[2] = 2 * ab;
    o[3] = o[2] + r(l1 * (uint64_t)char_width);
    o[4] = o[3] + r(l2 * 4);
    o[5] = o[4] + r(l1 * 4);
    o[6] = o[5] + r(F * 4);
    o[7] = o[6] + r((2 * F + 2) * 2 * 8);
    o[8] = o[7] + r(l1 * 4);
    o[9] = o[8] + r(l1 * 4);
}
__host__ __device__ inline uint64_t pbig_ws_bytes(int char_width, uint64_t l1, uint64_t l2)
{
    uint64_t o[10];
    pbig_layout(char_width, l1, l2, o);
    return o[9];
}
struct PbigArgs {
    const void* achars; // alignment: pairs (a[i], b[i]); cdist: queries (a) x choices (b)
    const uint64_t* aoff;
    const void* bchars;
    const uint64_t* boff;
    const uint32_t* items; // alignment: the pairs; cdist: the long queries (item = qk * nb + choice)
    uint64_t nitems;
    uint32_t nb;           // cdist: choices per query
    uint64_t out_stride;
    int cdist;             // 1: out[q * out_stride + c] = partial_ratio score (double)
    double cutoff;
    const double* pcut;    // cdist: per-pair cutoff (nullptr: cutoff)
    double* score_out;     // cdist output
    uint64_t* al_out;      // alignment output: 5 words per pair (rfa_score_alignment)
    uint8_t* ws;           // per-wave workspace of ws_stride bytes
    uint64_t ws_stride;
};
hipError_t launch_pbig(int char_width, const PbigArgs& a, uint32_t nblocks, hipStream_t s);

// Widest shorter string (in 64-unit words) the long path takes.
constexpr uint32_t PARTIAL_MAX_WORDS = 16;
hipError_t launch_plong(int char_width, uint32_t needle_words, const PartialLongArgs& a, uint32_t nblocks,
                        hipStream_t s);

size_t plan_temp_bytes(uint64_t n);
// Counting sort of the pairs into kernel ranges (items vals_out[k] = pair | pattern-is-B << 31).
// counters[0]: block queue; [1]: end of the block range; [2], [3]: ends of the OSA
// 33-64 and 1-32 unit ranges; [4 + r]: end of segmented range r (r = 0..13).
hipError_t launch_plan(const uint64_t* aoff, const uint64_t* boff, const uint8_t* metric, uint64_t n,
                       uint32_t chain_steps, uint32_t* bins, uint32_t* vals_in, uint32_t* vals_out, void* temp,
                       size_t temp_bytes, uint32_t* counters, hipStream_t s);

// Top-k scoring against a resident choice index (rfa_serve.hip).  Choices are
// sorted by length into groups of 64; group g holds gdwords[g] dwords per lane
// at gdata[gbase[g] + d * 64 + lane].  Queries [q_begin, q_end) of the batch
// are split into subsets of serve_queries_per_block(wq) (grid y); query q
// visits groups [qglo[q], qghi[q]), subset s the union [sub_glo[s], sub_ghi[s]).
struct ServeArgs {
    const uint32_t* gdata;
    const uint64_t* gbase;
    const uint32_t* gdwords;
    const uint32_t* glen_min;
    const uint32_t* slot_len;
    const uint32_t* slot_idx; // 0xffffffff: padding slot
    const uint8_t* qchars;
    const uint64_t* qoff;
    const uint32_t* qglo;
    const uint32_t* qghi;
    const uint32_t* sub_glo;
    const uint32_t* sub_ghi;
    uint32_t q_begin, q_end;
    ScoreSpec spec;
    double cutoff; // candidates: score >= cutoff
    uint32_t* counts;
    double* cand_score;
    int64_t* cand_index;
    uint32_t cap; // candidates kept per query (counts may exceed it)
    // dense mode (cdist of a few queries): every result is written to
    // dense_out[qrow[q] * dense_stride + choice] and no candidates are kept
    uint64_t* dense_out;
    const uint32_t* qrow;
    uint64_t dense_stride;
    const uint64_t* qcut; // dense mode: integer cutoff per query row (nullptr: spec.cutoff_u)
};

// Per-query selection of the k best (score desc, index asc) entries with
// score >= cutoff from a candidate list and/or a dense score row.
struct SelectArgs {
    const double* lscore; // list source: row q at q * lstride
    const int64_t* lindex; // entries with index < 0 are skipped
    const uint32_t* lcount; // entries per row (nullptr: lfixed)
    uint32_t lfixed, lcap;
    uint64_t lstride;
    const double* dscore; // dense source: row q at q * dstride, dn entries, index dbase + t
    uint64_t dstride, dn;
    int64_t dbase;
    double cutoff;
    uint32_t k;
    double* out_score; // [nq][k]; missing entries get score 0, index -1
    int64_t* out_index;
};

uint32_t serve_queries_per_block(uint32_t wq);
hipError_t launch_serve(int algo, uint32_t wq, const ServeArgs& a, uint32_t gx, uint32_t nsub, hipStream_t s);
hipError_t launch_select(const SelectArgs& a, uint32_t nq, hipStream_t s);

size_t lane_arena_words(uint32_t maxw);
hipError_t launch_lane(int char_width, int maxw, const LaneArgs& a, uint32_t nblocks, hipStream_t s);
hipError_t launch_lane_short(int char_width, int word_bits, const LaneArgs& a, uint32_t nblocks, hipStream_t s);

} // namespace rfa
