// rfa_tokens.hip -- token-sorted copies of a string set: every string split at
// whitespace, its tokens sorted, joined with single spaces
// (detail::sorted_split(...).join(), common_impl.hpp:161-179,
// SplittedSentenceView.hpp:56-90).  fuzz::token_sort_ratio and
// partial_token_sort_ratio score these copies with ratio / partial_ratio
// (fuzz_impl.hpp:296-368).
//
// Whitespace (common_impl.hpp:90-159): 0x09-0x0D, 0x1C-0x20 for every width;
// 16/32-bit units also 0x85, 0xA0, 0x1680, 0x2000-0x200A, 0x2028, 0x2029,
// 0x202F, 0x205F, 0x3000.  Tokens order by lexicographical comparison of their
// code units (Range.hpp:214-218): unsigned, or as signed 8-bit values when the
// caller's units are `char`.
//
// Three passes, one wave per string:
//   k_tok_count  -- output length (non-space units + tokens - 1) and token count;
//   (exclusive scans of both, hipcub)
//   k_tok_split  -- each token's (start, length), in order of appearance;
//   k_tok_write  -- each lane ranks its tokens against all tokens of the string
//                   (ties by position: equal tokens are identical) and sums the
//                   lengths of the tokens before it, which is its output offset;
//                   no lane waits for another.
// The deduplicated form (SplittedSentenceView::dedupe, token_set_ratio's sets)
// runs k_tok_dedup after the split: a token equal to an earlier one is dropped,
// and the output length and token count are those of the kept tokens; the write
// then also emits each kept token's (start, length, hash) in sorted order.
#include <hipcub/hipcub.hpp>

#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

template <typename CharT>
__device__ __forceinline__ bool tok_space(uint32_t c)
{
    if ((c >= 0x09 && c <= 0x0D) || (c >= 0x1C && c <= 0x20)) return true;
    if constexpr (sizeof(CharT) == 1) {
        return false;
    }
    else {
        return c == 0x85 || c == 0xA0 || c == 0x1680 || (c >= 0x2000 && c <= 0x200A) || c == 0x2028 ||
               c == 0x2029 || c == 0x202F || c == 0x205F || c == 0x3000;
    }
}

constexpr uint32_t TOK_THREADS = 256;

template <typename CharT>
__global__ __launch_bounds__(TOK_THREADS) void k_tok_count(const CharT* __restrict__ chars,
                                                           const uint64_t* __restrict__ off, uint64_t count,
                                                           uint64_t* __restrict__ out_len, uint64_t* __restrict__ ntok)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * TOK_THREADS + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * TOK_THREADS) >> 6;
    for (uint64_t s = wave0; s < count; s += nwaves) {
        const uint64_t o = off[s];
        const uint64_t n = off[s + 1] - o;
        uint64_t nonspace = 0, starts = 0;
        for (uint64_t i0 = 0; i0 < n; i0 += 64) {
            const uint64_t i = i0 + lane;
            bool sp = true, start = false;
            if (i < n) {
                sp = tok_space<CharT>((uint32_t)chars[o + i]);
                start = !sp && (i == 0 || tok_space<CharT>((uint32_t)chars[o + i - 1]));
            }
            nonspace += __popcll(ballot(i < n && !sp));
            starts += __popcll(ballot(start));
        }
        if (lane == 0) {
            out_len[s] = nonspace + (starts ? starts - 1 : 0);
            ntok[s] = starts;
        }
    }
}

template <typename CharT>
__global__ __launch_bounds__(TOK_THREADS) void k_tok_split(const CharT* __restrict__ chars,
                                                           const uint64_t* __restrict__ off, uint64_t count,
                                                           const uint64_t* __restrict__ tok_off,
                                                           uint32_t* __restrict__ tok_start,
                                                           uint32_t* __restrict__ tok_len)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * TOK_THREADS + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * TOK_THREADS) >> 6;
    for (uint64_t s = wave0; s < count; s += nwaves) {
        const uint64_t o = off[s];
        const uint64_t n = off[s + 1] - o;
        uint64_t t = tok_off[s];
        for (uint64_t i0 = 0; i0 < n; i0 += 64) {
            const uint64_t i = i0 + lane;
            bool start = false;
            if (i < n)
                start = !tok_space<CharT>((uint32_t)chars[o + i]) &&
                        (i == 0 || tok_space<CharT>((uint32_t)chars[o + i - 1]));
            const uint64_t mask = ballot(start);
            if (start) {
                uint64_t e = i + 1;
                while (e < n && !tok_space<CharT>((uint32_t)chars[o + e])) ++e;
                const uint64_t k = t + __popcll(mask & ((1ull << lane) - 1));
                tok_start[k] = (uint32_t)i;
                tok_len[k] = (uint32_t)(e - i);
            }
            t += __popcll(mask);
        }
    }
}

// 64-bit FNV-1a of a token's code units (equal tokens hash equal; users compare the
// units of tokens whose hashes match)
template <typename CharT>
__device__ __forceinline__ uint64_t tok_hash(const CharT* __restrict__ p, uint32_t n)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (uint32_t k = 0; k < n; ++k) {
        h ^= (uint64_t)(uint32_t)p[k];
        h *= 0x100000001b3ull;
    }
    return h;
}

// code unit as the comparison sees it
template <typename CharT, bool SIGNED8>
__device__ __forceinline__ int64_t tok_key(CharT c)
{
    if constexpr (SIGNED8 && sizeof(CharT) == 1)
        return (int64_t)(int8_t)c;
    else
        return (int64_t)c;
}

// -1, 0, 1: lexicographical order of two tokens
template <typename CharT, bool SIGNED8>
__device__ __forceinline__ int tok_cmp(const CharT* __restrict__ a, uint32_t la, const CharT* __restrict__ b,
                                       uint32_t lb)
{
    const uint32_t l = la < lb ? la : lb;
    for (uint32_t k = 0; k < l; ++k) {
        const int64_t x = tok_key<CharT, SIGNED8>(a[k]), y = tok_key<CharT, SIGNED8>(b[k]);
        if (x != y) return x < y ? -1 : 1;
    }
    return la == lb ? 0 : (la < lb ? -1 : 1);
}

// dup[t] = 1 for a token equal to an earlier token of its string; out_len / ntok of the kept tokens
template <typename CharT>
__global__ __launch_bounds__(TOK_THREADS) void k_tok_dedup(const CharT* __restrict__ chars,
                                                           const uint64_t* __restrict__ off, uint64_t count,
                                                           const uint64_t* __restrict__ tok_off,
                                                           const uint32_t* __restrict__ tok_start,
                                                           const uint32_t* __restrict__ tok_len,
                                                           uint8_t* __restrict__ dup, uint64_t* __restrict__ out_len,
                                                           uint64_t* __restrict__ ntok)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * TOK_THREADS + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * TOK_THREADS) >> 6;
    for (uint64_t s = wave0; s < count; s += nwaves) {
        const CharT* src = chars + off[s];
        const uint64_t t0 = tok_off[s], T = tok_off[s + 1] - t0;
        uint64_t units = 0, kept = 0;
        for (uint64_t a = lane; a < T; a += 64) {
            const uint32_t sa = tok_start[t0 + a], la = tok_len[t0 + a];
            bool d = false;
            for (uint64_t b = 0; b < a && !d; ++b) {
                const uint32_t lb = tok_len[t0 + b];
                if (lb != la) continue;
                const uint32_t sb = tok_start[t0 + b];
                uint32_t k = 0;
                while (k < la && src[sa + k] == src[sb + k]) ++k;
                d = k == la;
            }
            dup[t0 + a] = d ? 1 : 0;
            if (!d) {
                units += la;
                ++kept;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            units += __shfl_xor(units, o, 64);
            kept += __shfl_xor(kept, o, 64);
        }
        if (lane == 0) {
            out_len[s] = units + (kept ? kept - 1 : 0);
            ntok[s] = kept;
        }
    }
}

template <typename CharT, bool SIGNED8, bool DEDUP>
__global__ __launch_bounds__(TOK_THREADS) void k_tok_write(const CharT* __restrict__ chars,
                                                           const uint64_t* __restrict__ off, uint64_t count,
                                                           const uint64_t* __restrict__ tok_off,
                                                           const uint32_t* __restrict__ tok_start,
                                                           const uint32_t* __restrict__ tok_len,
                                                           const uint64_t* __restrict__ out_off,
                                                           CharT* __restrict__ out, const uint8_t* __restrict__ dup,
                                                           const uint64_t* __restrict__ set_off,
                                                           uint32_t* __restrict__ set_start,
                                                           uint32_t* __restrict__ set_len,
                                                           uint64_t* __restrict__ set_hash)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave0 = ((uint64_t)blockIdx.x * TOK_THREADS + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * TOK_THREADS) >> 6;
    for (uint64_t s = wave0; s < count; s += nwaves) {
        const CharT* src = chars + off[s];
        const uint64_t t0 = tok_off[s], T = tok_off[s + 1] - t0;
        CharT* dst = out + out_off[s];
        const uint64_t Tout = DEDUP ? set_off[s + 1] - set_off[s] : T;
        for (uint64_t a = lane; a < T; a += 64) {
            if (DEDUP && dup[t0 + a]) continue;
            const uint32_t sa = tok_start[t0 + a], la = tok_len[t0 + a];
            uint64_t rank = 0, pos = 0;
            for (uint64_t b = 0; b < T; ++b) {
                if (b == a || (DEDUP && dup[t0 + b])) continue;
                const uint32_t sb = tok_start[t0 + b], lb = tok_len[t0 + b];
                const int c = tok_cmp<CharT, SIGNED8>(src + sb, lb, src + sa, la);
                if (c < 0 || (c == 0 && b < a)) {
                    ++rank;
                    pos += (uint64_t)lb + 1;
                }
            }
            for (uint32_t k = 0; k < la; ++k) dst[pos + k] = src[sa + k];
            if (rank + 1 < Tout) dst[pos + la] = (CharT)0x20;
            if constexpr (DEDUP) {
                const uint64_t e = set_off[s] + rank;
                set_start[e] = (uint32_t)pos;
                set_len[e] = la;
                set_hash[e] = tok_hash<CharT>(src + sa, la);
            }
        }
    }
}

size_t tokens_scan_bytes(uint64_t count)
{
    size_t b = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                           (int)(count + 1));
    return b;
}

template <typename CharT>
static hipError_t run_tokens(const TokenArgs& a, hipStream_t s)
{
    const uint64_t waves_needed = a.count;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((waves_needed + TOK_THREADS / 64 - 1) / (TOK_THREADS / 64), (uint64_t)a.num_cus * 16));
    const CharT* in = (const CharT*)a.chars;
    hipLaunchKernelGGL(k_tok_count<CharT>, dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off, a.count, a.out_len,
                       a.ntok);
    size_t sb = a.scan_bytes;
    hipError_t e = hipSuccess;
    if (!a.dedup) {
        e = hipcub::DeviceScan::ExclusiveSum(a.scan_temp, sb, a.out_len, a.out_off, (int)(a.count + 1), s);
        if (e != hipSuccess) return e;
    }
    sb = a.scan_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(a.scan_temp, sb, a.ntok, a.tok_off, (int)(a.count + 1), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tok_split<CharT>, dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off, a.count, a.tok_off,
                       a.tok_start, a.tok_len);
    if (a.dedup) {
        // kept tokens: the output length and the token count again (ntok is free after its scan)
        hipLaunchKernelGGL(k_tok_dedup<CharT>, dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off, a.count, a.tok_off,
                           a.tok_start, a.tok_len, a.dup, a.out_len, a.ntok);
        sb = a.scan_bytes;
        e = hipcub::DeviceScan::ExclusiveSum(a.scan_temp, sb, a.out_len, a.out_off, (int)(a.count + 1), s);
        if (e != hipSuccess) return e;
        sb = a.scan_bytes;
        e = hipcub::DeviceScan::ExclusiveSum(a.scan_temp, sb, a.ntok, a.set_off, (int)(a.count + 1), s);
        if (e != hipSuccess) return e;
        if (a.signed8)
            hipLaunchKernelGGL((k_tok_write<CharT, true, true>), dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off,
                               a.count, a.tok_off, a.tok_start, a.tok_len, a.out_off, (CharT*)a.out, a.dup, a.set_off,
                               a.set_start, a.set_len, a.set_hash);
        else
            hipLaunchKernelGGL((k_tok_write<CharT, false, true>), dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off,
                               a.count, a.tok_off, a.tok_start, a.tok_len, a.out_off, (CharT*)a.out, a.dup, a.set_off,
                               a.set_start, a.set_len, a.set_hash);
    }
    else if (a.signed8)
        hipLaunchKernelGGL((k_tok_write<CharT, true, false>), dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off,
                           a.count, a.tok_off, a.tok_start, a.tok_len, a.out_off, (CharT*)a.out, nullptr, nullptr,
                           nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_tok_write<CharT, false, false>), dim3(blocks), dim3(TOK_THREADS), 0, s, in, a.off,
                           a.count, a.tok_off, a.tok_start, a.tok_len, a.out_off, (CharT*)a.out, nullptr, nullptr,
                           nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_tokens(int char_width, const TokenArgs& a, hipStream_t s)
{
    if (char_width == 1) return run_tokens<uint8_t>(a, s);
    if (char_width == 2) return run_tokens<uint16_t>(a, s);
    return run_tokens<uint32_t>(a, s);
}

} // namespace rfa
