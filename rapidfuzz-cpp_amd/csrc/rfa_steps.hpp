// rfa_steps.hpp -- single-word bit-parallel recurrences shared by the query-tile
// kernel and the short-pair lane kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rfa {

template <typename W>
__device__ __forceinline__ int popc(W x)
{
    if constexpr (sizeof(W) == 8)
        return __popcll(x);
    else
        return __popc(x);
}

// ---- recurrences on one word (one query per lane) -----------------------
//
// The distance is not tracked per text position: after the last position the
// vertical deltas of the final column give D[m][n] = n + popcount(VP & M) -
// popcount(VN & M) with M the m low bits, so the loop body is the bare
// recurrence.  gfx950's three-input bitop (v_bitop3_b32, truth table over
// a=0xF0, b=0xCC, c=0xAA) folds the NOR/OR/XOR pairs, and 64-bit words are
// kept as two 32-bit halves (v_alignbit_b32 carries the shifted-out bit).

constexpr uint32_t TA = 0xF0, TB = 0xCC, TC = 0xAA;
constexpr uint32_t T_XOR_OR = ((TA ^ TB) | TC) & 0xFF;  // (a ^ b) | c
constexpr uint32_t T_OR_NOR = (TA | ~(TB | TC)) & 0xFF; // a | ~(b | c)
constexpr uint32_t T_OR_ANDN = (TA | (TB & ~TC)) & 0xFF; // a | (b & ~c)
constexpr uint32_t T_AND2 = (TA & TB) & 0xFF;             // a & b (c unused)

template <uint32_t IMM>
__device__ __forceinline__ uint32_t bop3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, IMM);
}

// (hi:lo) << 1, high half
__device__ __forceinline__ uint32_t shl1_hi(uint32_t hi, uint32_t lo) { return __builtin_amdgcn_alignbit(hi, lo, 31); }

// Levenshtein_impl.hpp:282-309 (score handled at the end)
__device__ __forceinline__ void lev_step(uint32_t X, uint32_t& VP, uint32_t& VN)
{
    const uint32_t XV = X | VN;
    const uint32_t s = (X & VP) + VP;
    const uint32_t D0 = bop3<T_XOR_OR>(s, VP, XV);
    uint32_t HP = bop3<T_OR_NOR>(VN, D0, VP);
    uint32_t HN = D0 & VP;
    HP = (HP << 1) | 1; // one v_lshl_or_b32: VN below reads it through bitop3, which the
    HN = HN << 1;       // compiler does not fold back into a separate shift and OR
    VP = bop3<T_OR_NOR>(HN, D0, HP);
    VN = bop3<T_AND2>(HP, D0, 0u);
}

__device__ __forceinline__ void lev_step(uint64_t X, uint32_t& VPl, uint32_t& VPh, uint32_t& VNl, uint32_t& VNh)
{
    const uint32_t Xl = (uint32_t)X, Xh = (uint32_t)(X >> 32);
    const uint32_t XVl = Xl | VNl, XVh = Xh | VNh;
    const uint64_t t = ((uint64_t)(Xh & VPh) << 32) | (Xl & VPl);
    const uint64_t s = t + (((uint64_t)VPh << 32) | VPl);
    const uint32_t D0l = bop3<T_XOR_OR>((uint32_t)s, VPl, XVl);
    const uint32_t D0h = bop3<T_XOR_OR>((uint32_t)(s >> 32), VPh, XVh);
    const uint32_t HPl0 = bop3<T_OR_NOR>(VNl, D0l, VPl);
    const uint32_t HPh0 = bop3<T_OR_NOR>(VNh, D0h, VPh);
    const uint32_t HNl0 = D0l & VPl, HNh0 = D0h & VPh;
    const uint32_t HPh = shl1_hi(HPh0, HPl0), HPl = (HPl0 << 1) | 1;
    const uint32_t HNh = shl1_hi(HNh0, HNl0), HNl = HNl0 << 1;
    VPl = bop3<T_OR_NOR>(HNl, D0l, HPl);
    VPh = bop3<T_OR_NOR>(HNh, D0h, HPh);
    VNl = bop3<T_AND2>(HPl, D0l, 0u);
    VNh = HPh & D0h;
}

// LCSseq_impl.hpp:219-252 (one word): S = (S + u) | (S - u), S - u = S & ~X
__device__ __forceinline__ void lcs_step(uint32_t X, uint32_t& S)
{
    const uint32_t u = S & X;
    S = bop3<T_OR_ANDN>(S + u, S, X);
}

__device__ __forceinline__ void lcs_step(uint64_t X, uint32_t& Sl, uint32_t& Sh)
{
    const uint32_t Xl = (uint32_t)X, Xh = (uint32_t)(X >> 32);
    const uint64_t S = ((uint64_t)Sh << 32) | Sl;
    const uint64_t sum = S + (((uint64_t)(Sh & Xh) << 32) | (Sl & Xl));
    Sl = bop3<T_OR_ANDN>((uint32_t)sum, Sl, Xl);
    Sh = bop3<T_OR_ANDN>((uint32_t)(sum >> 32), Sh, Xh);
}

// OSA_impl.hpp:48-71 (score handled at the end)
__device__ __forceinline__ void osa_step(uint32_t X, uint32_t& VP, uint32_t& VN, uint32_t& D0, uint32_t& PMold)
{
    const uint32_t TR = ((X & ~D0) << 1) & PMold;
    const uint32_t XVT = X | VN | TR;
    const uint32_t s = (X & VP) + VP;
    D0 = bop3<T_XOR_OR>(s, VP, XVT);
    uint32_t HP = bop3<T_OR_NOR>(VN, D0, VP);
    uint32_t HN = D0 & VP;
    HP = (HP << 1) | 1;
    HN = HN << 1;
    VP = bop3<T_OR_NOR>(HN, D0, HP);
    VN = bop3<T_AND2>(HP, D0, 0u);
    PMold = X;
}

__device__ __forceinline__ void osa_step(uint64_t X, uint32_t& VPl, uint32_t& VPh, uint32_t& VNl, uint32_t& VNh,
                                         uint32_t& D0l, uint32_t& D0h, uint32_t& PMl, uint32_t& PMh)
{
    const uint32_t Xl = (uint32_t)X, Xh = (uint32_t)(X >> 32);
    const uint32_t al = Xl & ~D0l, ah = Xh & ~D0h;
    const uint32_t TRl = (al << 1) & PMl, TRh = shl1_hi(ah, al) & PMh;
    const uint32_t XVTl = Xl | VNl | TRl, XVTh = Xh | VNh | TRh;
    const uint64_t t = ((uint64_t)(Xh & VPh) << 32) | (Xl & VPl);
    const uint64_t s = t + (((uint64_t)VPh << 32) | VPl);
    D0l = bop3<T_XOR_OR>((uint32_t)s, VPl, XVTl);
    D0h = bop3<T_XOR_OR>((uint32_t)(s >> 32), VPh, XVTh);
    const uint32_t HPl0 = bop3<T_OR_NOR>(VNl, D0l, VPl);
    const uint32_t HPh0 = bop3<T_OR_NOR>(VNh, D0h, VPh);
    const uint32_t HNl0 = D0l & VPl, HNh0 = D0h & VPh;
    const uint32_t HPh = shl1_hi(HPh0, HPl0), HPl = (HPl0 << 1) | 1;
    const uint32_t HNh = shl1_hi(HNh0, HNl0), HNl = HNl0 << 1;
    VPl = bop3<T_OR_NOR>(HNl, D0l, HPl);
    VPh = bop3<T_OR_NOR>(HNh, D0h, HPh);
    VNl = HPl & D0l;
    VNh = HPh & D0h;
    PMl = Xl;
    PMh = Xh;
}

} // namespace rfa
