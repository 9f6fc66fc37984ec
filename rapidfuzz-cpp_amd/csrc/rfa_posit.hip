// rfa_posit.hip -- positional metrics for cdist: Hamming with padding
// (Hamming_impl.hpp:23-35: max(m, n) minus the equal positions of the common
// length), common prefix and common suffix length (Prefix_impl.hpp:23-27,
// Postfix_impl.hpp:23-27).
//
// A workgroup holds a tile of 64 queries, one per lane, in each of its 4 waves;
// the waves take runs of 8 consecutive choices.  A choice's code units are
// wave-uniform (one broadcast load per 8 units), each lane compares them with
// its own query and keeps its 8 results in registers, stored as one contiguous
// 64-byte row segment.  No bit-parallel recurrence: the work is one compare
// per unit, so the kernel is bound by the result stores.
#include <algorithm>

#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"

namespace rfa {

constexpr uint32_t POSIT_RUN = 8;

template <int METRIC, typename CharT, bool INTK>
__global__ __launch_bounds__(256) void k_posit(PositArgs a)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t q = a.q_base + blockIdx.y * 64 + lane;
    const bool qv = q < a.nq;
    const CharT* __restrict__ QC = (const CharT*)a.qchars;
    const CharT* __restrict__ TC = (const CharT*)a.cchars;
    const uint64_t qo = qv ? a.qoff[q] : 0;
    const uint32_t m = qv ? (uint32_t)(a.qoff[q + 1] - qo) : 0u;
    const CharT* Q = QC + qo;
    uint32_t mmax = m;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mmax = max(mmax, (uint32_t)__shfl_xor(mmax, off, 64));
    mmax = uniform(mmax);
    ScoreSpec spec = a.spec;
    if (a.qcut && qv) spec.cutoff_u = a.qcut[q];
    const uint32_t nruns = (a.nc + POSIT_RUN - 1) / POSIT_RUN;
    for (uint32_t run = blockIdx.x * 4 + wave; run < nruns; run += gridDim.x * 4) {
        const uint32_t c0 = run * POSIT_RUN;
        uint64_t res[POSIT_RUN];
#pragma unroll
        for (uint32_t k = 0; k < POSIT_RUN; ++k) {
            res[k] = 0;
            const uint32_t c = c0 + k;
            if (c >= a.nc) break;
            const uint64_t to = a.coff[c];
            const uint32_t n = (uint32_t)(a.coff[c + 1] - to);
            const CharT* T = TC + to;
            const uint32_t both = min(m, n);
            const uint32_t lim = min(n, mmax); // wave-uniform
            uint64_t raw;
            if constexpr (METRIC == RFA_POSTFIX) {
                uint32_t len = 0;
                bool alive = qv;
                for (uint32_t i = 0; i < lim; ++i) {
                    const uint32_t tu = (uint32_t)T[n - 1 - i];
                    const bool ok = alive && i < both && (uint32_t)Q[m - 1 - (i < m ? i : 0)] == tu;
                    len += ok;
                    alive = ok;
                    if (!ballot(alive)) break;
                }
                raw = len;
            }
            else {
                // 8 units at a time; a lane past its query's end re-reads the query's start
                // (the 64-byte tail padding of the set covers the short over-read)
                uint32_t cnt = 0;
                bool alive = qv;
                for (uint32_t p0 = 0; p0 < lim; p0 += 8) {
                    uint32_t tu[8], qu[8];
                    load_units<CharT, 8>(T + p0, tu);
                    load_units<CharT, 8>(p0 < m ? Q + p0 : Q, qu);
#pragma unroll
                    for (uint32_t j = 0; j < 8; ++j) {
                        const bool eq = p0 + j < both && tu[j] == qu[j];
                        if constexpr (METRIC == RFA_HAMMING) {
                            cnt += eq;
                        }
                        else {
                            alive = alive && eq;
                            cnt += alive;
                        }
                    }
                    if constexpr (METRIC == RFA_PREFIX) {
                        if (!ballot(alive)) break;
                    }
                }
                raw = METRIC == RFA_HAMMING ? (uint64_t)max(m, n) - cnt : cnt;
            }
            if constexpr (INTK) {
                res[k] = finalize_int(spec, m, n, raw);
            }
            else {
                RawResult r{raw, 0, 0};
                res[k] = finalize_bits(spec, m, n, r);
            }
        }
        if (qv) {
            uint64_t* row = (uint64_t*)a.out + (uint64_t)q * a.out_stride + c0;
            const uint32_t cnt = min(POSIT_RUN, a.nc - c0);
#pragma unroll
            for (uint32_t k = 0; k < POSIT_RUN; ++k)
                if (k < cnt) row[k] = res[k];
        }
    }
}

template <int METRIC, typename CharT>
static void launch_posit_m(const PositArgs& a, uint32_t gx, uint32_t gy, hipStream_t s)
{
    if (a.spec.kind == RFA_DISTANCE || a.spec.kind == RFA_SIMILARITY)
        hipLaunchKernelGGL((k_posit<METRIC, CharT, true>), dim3(gx, gy), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_posit<METRIC, CharT, false>), dim3(gx, gy), dim3(256), 0, s, a);
}

template <typename CharT>
static void launch_posit_w(const PositArgs& a, uint32_t gx, uint32_t gy, hipStream_t s)
{
    switch (a.spec.metric) {
    case RFA_HAMMING: launch_posit_m<RFA_HAMMING, CharT>(a, gx, gy, s); break;
    case RFA_PREFIX: launch_posit_m<RFA_PREFIX, CharT>(a, gx, gy, s); break;
    default: launch_posit_m<RFA_POSTFIX, CharT>(a, gx, gy, s); break;
    }
}

hipError_t launch_posit(int char_width, const PositArgs& a, uint32_t num_cus, hipStream_t s)
{
    const uint32_t tiles = (a.nq + 63) / 64;
    const uint32_t nruns = (a.nc + POSIT_RUN - 1) / POSIT_RUN;
    // grid y: query tiles, at most 65,535 per launch
    for (uint32_t t0 = 0; t0 < tiles; t0 += 65535) {
        PositArgs b = a;
        b.q_base = t0 * 64;
        const uint32_t gy = std::min<uint32_t>(65535, tiles - t0);
        // about 32 waves per CU in total, no more workgroups than runs
        const uint64_t want = ((uint64_t)num_cus * 32 / 4 + gy - 1) / gy;
        const uint32_t gx = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (nruns + 3) / 4));
        if (char_width == 1)
            launch_posit_w<uint8_t>(b, gx, gy, s);
        else if (char_width == 2)
            launch_posit_w<uint16_t>(b, gx, gy, s);
        else
            launch_posit_w<uint32_t>(b, gx, gy, s);
    }
    return hipGetLastError();
}

} // namespace rfa
