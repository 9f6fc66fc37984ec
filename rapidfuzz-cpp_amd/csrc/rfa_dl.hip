// rfa_dl.hip -- experimental::CachedDamerauLevenshtein for cdist: the unrestricted
// Damerau-Levenshtein distance (insertions, deletions, substitutions and
// transpositions of adjacent units with edits between them), which the reference
// computes with Zhao and Sahni's linear-space algorithm
// (DamerauLevenshtein_impl.hpp:36-121) and cuts like the other distances.
//
// The distance is symmetric in its two strings, so the kernel runs the same
// algorithm with the roles of the reference's loops exchanged: the choice is the
// outer string (one column c per choice unit, uniform across the wave) and each
// lane's query is the inner one (rows r = 1..m in registers).  Zhao's state maps
// onto registers per lane:
//   A, B   the columns c - 2 (overwritten with column c) and c - 1;
//   G[r]   FR[r] - k: the value saved at the last match of row r (column k - 1,
//          row r - 2) minus that match's column k, so the transposition through
//          it costs G[r] + c;
//   U      T - l: the value saved at the last match of the current column (row l,
//          column c - 2, row l - 1) minus l, so the transposition through it
//          costs U + r.
// Zhao uses the first term when row r - 1 matched the current unit and the second
// when the previous unit matched row r (it tries the first one first).  Both are
// the Lowrance-Wagner term D(k-1, l-1) + (r-k-1) + 1 + (c-l-1) for the last
// occurrences k, l, and that term never falls below the distance, so the kernel
// takes the minimum of whichever apply, and applies them on matches too, where
// they cannot win.  The reference's sentinel max(|s1|, |s2|) + 1 becomes a larger
// constant: it only ever enters terms that lose.
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_tile_pm.hpp"

namespace rfa {

namespace {

constexpr uint32_t DL_BIG = 0x40000000u;

__device__ __forceinline__ uint32_t dl_wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor(v, off, 64));
    return uniform(v);
}

template <uint32_t I>
__device__ __forceinline__ uint32_t half_of(uint32_t lo, uint32_t hi)
{
    return I < 32 ? lo : hi;
}

// (m & a) | (~m & b), written as the instruction: the compiler would otherwise turn a
// 0 / -1 mask back into a compare and a select, whose masks it then spills.
__device__ __forceinline__ uint32_t bitsel(uint32_t m, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}

// Rows I..BITS of column c, written into A over column c - 2 (B holds column c - 1).
// (xlo, xhi): this unit's match bits; (ylo, yhi) their complement; (qlo, qhi): the
// complement of the previous unit's; li = column c - 2's value of row I - 1 before
// it was overwritten.  Every choice is arithmetic, so no per-row condition masks
// are kept: the match bit as 0 / -1 (v_bfe_i32), a term that does not apply
// carries the sentinel bit (v_lshl_or_b32), the saved values update under the
// match mask (v_bfi_b32).  Real values stay below 2^29, so a term with bit 30 set
// never wins.
template <uint32_t BITS, uint32_t I = 1>
__device__ __forceinline__ void dl_rows(uint32_t (&A)[BITS + 1], const uint32_t (&B)[BITS + 1],
                                        uint32_t (&G)[BITS + 1], uint32_t xlo, uint32_t xhi, uint32_t ylo,
                                        uint32_t yhi, uint32_t qlo, uint32_t qhi, uint32_t c, uint32_t mt,
                                        uint32_t li, uint32_t U)
{
    if constexpr (I <= BITS) {
        if ((I & 3) == 1 && I > mt) return;
        const uint32_t old = A[I];
        const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)half_of<I - 1>(xlo, xhi), (I - 1) & 31, 1);
        const uint32_t miss = m + 1u;
        uint32_t t = min(min(B[I - 1] + miss, A[I - 1] + 1u), B[I] + 1u);
        // row I - 1 matched this unit: the transposition through G[I]
        const uint32_t n1 = I >= 2 ? __builtin_amdgcn_ubfe(half_of<(I >= 2 ? I - 2 : 0)>(ylo, yhi), (I - 2) & 31, 1)
                                   : 1u;
        const uint32_t t1 = (G[I] + c) | (n1 << 30);
        // the previous unit matched row I: the transposition through U
        const uint32_t n2 = __builtin_amdgcn_ubfe(half_of<I - 1>(qlo, qhi), (I - 1) & 31, 1);
        const uint32_t t2 = (U + I) | (n2 << 30);
        t = min(t, min(t1, t2));
        const uint32_t bm2 = I >= 2 ? B[I >= 2 ? I - 2 : 0] : DL_BIG;
        G[I] = bitsel(m, bm2 - c, G[I]);
        U = bitsel(m, li - I, U);
        A[I] = t;
        dl_rows<BITS, I + 1>(A, B, G, xlo, xhi, ylo, yhi, qlo, qhi, c, mt, old, U);
    }
}

template <typename W, uint32_t BITS>
__device__ __forceinline__ void dl_step(uint32_t (&A)[BITS + 1], const uint32_t (&B)[BITS + 1],
                                        uint32_t (&G)[BITS + 1], W X, W P, uint32_t c, uint32_t mt)
{
    const uint32_t xlo = (uint32_t)X;
    const uint32_t xhi = BITS > 32 ? (uint32_t)((uint64_t)X >> 32) : 0u;
    const uint32_t qlo = ~(uint32_t)P;
    const uint32_t qhi = BITS > 32 ? ~(uint32_t)((uint64_t)P >> 32) : ~0u;
    const uint32_t li = A[0];
    A[0] = c;
    dl_rows<BITS>(A, B, G, xlo, xhi, ~xlo, ~xhi, qlo, qhi, c, mt, li, DL_BIG);
}

} // namespace

template <typename CharT, typename W, bool INTK>
__global__ __launch_bounds__(TILE_THREADS) void k_dltile(TileArgs a)
{
    extern __shared__ uint64_t smem_u64[];
    W* pm = (W*)smem_u64;
    uint8_t* extra = (uint8_t*)(pm + (size_t)a.rows_cap * 64);
    const CharT* __restrict__ q = (const CharT*)a.qchars;
    const CharT* __restrict__ c = (const CharT*)a.cchars;
    constexpr uint32_t BITS = sizeof(W) * 8;

    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t tile = blockIdx.y + a.tile_base;
    const uint32_t qpos = tile * 64 + lane;
    const bool qact = qpos < a.nq;
    const uint32_t qidx = qact ? a.qorder[qpos] : 0;
    const uint64_t qo = qact ? a.qoff[qidx] : 0;
    const uint32_t m = qact ? (uint32_t)(a.qoff[qidx + 1] - qo) : 0;

    SymLookup<CharT> look = build_tile_pm<CharT, W>(q, qo, m, pm, extra, a);
    ScoreSpec spec = a.spec;
    if (a.qcut && qact) spec.cutoff_u = a.qcut[qidx];
    const uint32_t mt = dl_wave_max(m);
    const W* __restrict__ col = pm + lane;

    const uint32_t it0 = blockIdx.x * a.choices_per_block;
    const uint32_t it1 = min(a.ncsel, it0 + a.choices_per_block);
    constexpr uint32_t GR = 4;
    for (uint32_t g0 = it0 + uniform(wave) * GR; g0 < it1; g0 += TILE_WAVES * GR) {
        const uint32_t gn = min(GR, it1 - g0);
        uint64_t outv[GR];
        for (uint32_t t = 0; t < gn; ++t) {
            const uint32_t cidx = g0 + t;
            const uint64_t co0 = a.coff[cidx];
            const uint32_t n = uniform((uint32_t)(a.coff[cidx + 1] - co0));
            const uint64_t co = ((uint64_t)uniform((uint32_t)(co0 >> 32)) << 32) | uniform((uint32_t)co0);
            // A: column -1 (all sentinel), B: column 0 (D(0, r) = r)
            uint32_t A[BITS + 1], B[BITS + 1], G[BITS + 1];
#pragma unroll
            for (uint32_t i = 0; i <= BITS; ++i) {
                A[i] = DL_BIG;
                B[i] = i;
                G[i] = DL_BIG;
            }
            W P = 0;
            uint32_t unit_next = lane < n ? (uint32_t)c[co + lane] : 0u;
            for (uint32_t k0 = 0; k0 < n; k0 += 64) {
                const uint32_t row = k0 + lane < n ? look(unit_next) : 0u;
                if (k0 + 64 + lane < n) unit_next = (uint32_t)c[co + k0 + 64 + lane];
                const uint32_t steps = min(64u, n - k0);
                W Xn = col[readlane(row, 0) * 64];
                // columns alternate between A (odd c) and B (even c); k0 is even
                for (uint32_t j = 0; j < steps; j += 2) {
                    const W X1 = Xn;
                    Xn = col[readlane(row, min(j + 1, 63u)) * 64];
                    dl_step<W, BITS>(A, B, G, X1, P, k0 + j + 1, mt);
                    P = X1;
                    if (j + 1 < steps) {
                        const W X2 = Xn;
                        Xn = col[readlane(row, min(j + 2, 63u)) * 64];
                        dl_step<W, BITS>(B, A, G, X2, P, k0 + j + 2, mt);
                        P = X2;
                    }
                }
            }
            uint32_t res = (n & 1) ? A[0] : B[0];
#pragma unroll
            for (uint32_t i = 1; i <= BITS; ++i) res = (i == m) ? ((n & 1) ? A[i] : B[i]) : res;
            RawResult r{};
            r.v = res;
            if constexpr (INTK)
                outv[t] = finalize_int(spec, m, n, res);
            else
                outv[t] = finalize_bits(spec, m, n, r);
        }
        if (qact) {
            uint64_t* dst = (uint64_t*)a.out + (uint64_t)qidx * a.out_stride + g0;
            if (gn == GR && ((uintptr_t)dst & 15) == 0) {
                uint4* d4 = (uint4*)dst;
#pragma unroll
                for (uint32_t t = 0; t < GR; t += 2)
                    d4[t / 2] = make_uint4((uint32_t)outv[t], (uint32_t)(outv[t] >> 32), (uint32_t)outv[t + 1],
                                           (uint32_t)(outv[t + 1] >> 32));
            }
            else {
                for (uint32_t t = 0; t < gn; ++t) dst[t] = outv[t];
            }
        }
    }
}

// ---- long queries: one lane per (query, choice) pair ---------------------------
//
// A wave takes one query and 64 consecutive choices.  The query's units and the row
// loop are uniform across the wave; each lane streams its own choice.  The three
// arrays of the tile kernel (two columns and G) live in the wave's slice of global
// scratch, lane-interleaved ([row][lane], so every access of the wave is one 256-byte
// segment); the cell is the tile kernel's, with the match tests as compares.
template <typename CharT, bool INTK>
__global__ __launch_bounds__(64) void k_dllong(DlLongArgs a)
{
    const CharT* __restrict__ Q = (const CharT*)a.qchars;
    const CharT* __restrict__ C = (const CharT*)a.cchars;
    const uint32_t lane = threadIdx.x;
    const uint64_t stride = (uint64_t)(a.max_rows + 1) * 64;
    uint32_t* __restrict__ S0 = a.scratch + (uint64_t)blockIdx.x * 3 * stride + lane;
    uint32_t* __restrict__ S1 = S0 + stride;
    uint32_t* __restrict__ Gs = S1 + stride;
    const uint64_t nchunks = (a.nc + 63) / 64;
    const uint64_t items = (uint64_t)a.nql * nchunks;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t qi = a.qlist[it / nchunks];
        const uint64_t cj = (it % nchunks) * 64 + lane;
        const bool active = cj < a.nc;
        const uint64_t qo = a.qoff[qi];
        const uint32_t m = uniform((uint32_t)(a.qoff[qi + 1] - qo));
        const uint64_t co = active ? a.coff[cj] : 0;
        const uint32_t n = active ? (uint32_t)(a.coff[cj + 1] - co) : 0u;
        uint32_t nmax = n;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) nmax = max(nmax, (uint32_t)__shfl_xor(nmax, off, 64));
        nmax = uniform(nmax);
        // S0: column -1 (all sentinel), S1: column 0 (D(0, r) = r)
        for (uint32_t r = 0; r <= m; ++r) {
            S0[(uint64_t)r * 64] = DL_BIG;
            S1[(uint64_t)r * 64] = r;
            Gs[(uint64_t)r * 64] = DL_BIG;
        }
        uint32_t prev_ch = 0;
        bool have_prev = false;
        for (uint32_t c = 1; c <= nmax; ++c) {
            const bool act = c <= n;
            const uint32_t ch = act ? (uint32_t)C[co + c - 1] : 0u;
            // odd columns overwrite S0 (column c - 2), even ones S1
            uint32_t* __restrict__ X = (c & 1) ? S0 : S1;
            const uint32_t* __restrict__ Y = (c & 1) ? S1 : S0;
            if (act) {
                uint32_t li = X[0];
                X[0] = c;
                uint32_t U = DL_BIG;
                uint32_t yprev = Y[0], xprev = c, bm2 = DL_BIG;
                uint32_t qprev = 0;
                for (uint32_t r = 1; r <= m; ++r) {
                    const uint64_t o = (uint64_t)r * 64;
                    const uint32_t old = X[o], yr = Y[o], g = Gs[o];
                    const uint32_t qr = (uint32_t)Q[qo + r - 1];
                    const bool match = qr == ch;
                    uint32_t t = min(min(yprev + (match ? 0u : 1u), xprev + 1u), yr + 1u);
                    if (r >= 2 && qprev == ch) t = min(t, g + c);             // row r - 1 matched this unit
                    if (have_prev && qr == prev_ch) t = min(t, U + r);       // the previous unit matched row r
                    if (match) {
                        Gs[o] = bm2 - c;
                        U = li - r;
                    }
                    X[o] = t;
                    li = old;
                    bm2 = yprev;
                    yprev = yr;
                    xprev = t;
                    qprev = qr;
                }
            }
            prev_ch = ch;
            have_prev = true;
        }
        if (active) {
            const uint32_t res = ((n & 1) ? S0 : S1)[(uint64_t)m * 64];
            ScoreSpec spec = a.spec;
            if (a.qcut) spec.cutoff_u = a.qcut[qi];
            RawResult rr{};
            rr.v = res;
            const uint64_t bits = INTK ? finalize_int(spec, m, n, res) : finalize_bits(spec, m, n, rr);
            ((uint64_t*)a.out)[(uint64_t)qi * a.out_stride + cj] = bits;
        }
    }
}

template <typename CharT>
static void launch_dllong_c(const DlLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY)
        hipLaunchKernelGGL((k_dllong<CharT, true>), dim3(nblocks), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_dllong<CharT, false>), dim3(nblocks), dim3(64), 0, s, a);
}

hipError_t launch_dllong(int char_width, const DlLongArgs& a, uint32_t nblocks, hipStream_t s)
{
    if (char_width == 1)
        launch_dllong_c<uint8_t>(a, nblocks, s);
    else if (char_width == 2)
        launch_dllong_c<uint16_t>(a, nblocks, s);
    else
        launch_dllong_c<uint32_t>(a, nblocks, s);
    return hipGetLastError();
}

template <typename CharT, typename W>
static void launch_dltile_w(const TileArgs& a, uint32_t ntiles, uint32_t nbx, size_t lds, hipStream_t s)
{
    const bool intk = a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY;
    if (intk)
        hipLaunchKernelGGL((k_dltile<CharT, W, true>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a);
    else
        hipLaunchKernelGGL((k_dltile<CharT, W, false>), dim3(nbx, ntiles), dim3(TILE_THREADS), lds, s, a);
}

hipError_t launch_dltile(int char_width, int word_bits, const TileArgs& a, uint32_t ntiles, uint32_t nblocks_x,
                         hipStream_t s)
{
    const size_t lds = tile_lds_bytes(char_width, word_bits, a.rows_cap, a.hash_bits);
    if (word_bits == 32) {
        if (char_width == 1)
            launch_dltile_w<uint8_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_dltile_w<uint16_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_dltile_w<uint32_t, uint32_t>(a, ntiles, nblocks_x, lds, s);
    }
    else {
        if (char_width == 1)
            launch_dltile_w<uint8_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else if (char_width == 2)
            launch_dltile_w<uint16_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
        else
            launch_dltile_w<uint32_t, uint64_t>(a, ntiles, nblocks_x, lds, s);
    }
    return hipGetLastError();
}

} // namespace rfa
