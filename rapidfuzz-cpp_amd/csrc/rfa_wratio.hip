// rfa_wratio.hip -- fuzz::WRatio's combining steps (fuzz_impl.hpp:790-826,
// CachedWRatio::similarity fuzz_impl.hpp:838-876) over a cdist matrix.
//
// WRatio scores a pair with ratio first (end_ratio), then, by the length ratio
// lr = max(len) / min(len):
//   lr < 1.5:  max(end_ratio, token_ratio(cutoff') * 0.95),
//              cutoff' = max(cutoff, end_ratio) / 0.95;
//   lr >= 1.5: with PS = 0.9 (lr <= 8) or 0.6,
//              e = max(end_ratio, partial_ratio(c2) * PS), c2 = max(cutoff, end_ratio) / PS;
//              max(e, partial_token_ratio(c3) * 0.95 * PS), c3 = max(c2, e) / 0.95.
// The sub-scores are whole cdist launches that read these raised cutoffs per
// pair; the steps here compute the cutoffs and the final maximum in the
// reference's order of double operations.  An empty string scores 0.
#include "rfa_device.hpp"
#include "rfa_internal.hpp"

namespace rfa {

__device__ __forceinline__ double wr_max(double a, double b) { return a < b ? b : a; } // std::max(a, b)

// STAGE 0: R = end_ratio -> C = the first raised cutoff.
// STAGE 1 (lr >= 1.5): X = partial_ratio(C) -> R = e, C = c3.
// STAGE 2: X = partial_token_ratio(C) (lr >= 1.5), out = token_ratio(C) (lr < 1.5) -> out = WRatio.
template <int STAGE>
__global__ __launch_bounds__(256) void k_wratio(WRatioArgs a)
{
    const uint64_t total = (uint64_t)a.nq * a.nc;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(i / a.nc), c = (uint32_t)(i - (uint64_t)q * a.nc);
        const uint64_t len1 = a.qoff[q + 1] - a.qoff[q], len2 = a.coff[c + 1] - a.coff[c];
        if (len1 == 0 || len2 == 0) {
            if constexpr (STAGE == 0) a.C[i] = 101.0; // every sub-score is 0 above 100
            if constexpr (STAGE == 2) a.out[i] = 0.0;
            continue;
        }
        const double lr = len1 > len2 ? (double)len1 / (double)len2 : (double)len2 / (double)len1;
        const double ps = lr <= 8.0 ? 0.9 : 0.6;
        if constexpr (STAGE == 0) {
            const double m = wr_max(a.cutoff, a.R[i]);
            a.C[i] = lr < 1.5 ? m / 0.95 : m / ps;
        }
        else if constexpr (STAGE == 1) {
            if (lr >= 1.5) {
                const double e = wr_max(a.R[i], a.X[i] * ps);
                a.R[i] = e;
                a.C[i] = wr_max(a.C[i], e) / 0.95;
            }
        }
        else {
            if (lr < 1.5)
                a.out[i] = wr_max(a.R[i], a.out[i] * 0.95);
            else
                a.out[i] = wr_max(a.R[i], a.X[i] * 0.95 * ps);
        }
    }
}

hipError_t launch_wratio(int stage, const WRatioArgs& a, uint32_t num_cus, hipStream_t s)
{
    const uint64_t total = (uint64_t)a.nq * a.nc;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, (uint64_t)num_cus * 16));
    switch (stage) {
    case 0: hipLaunchKernelGGL((k_wratio<0>), dim3(blocks), dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_wratio<1>), dim3(blocks), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_wratio<2>), dim3(blocks), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

} // namespace rfa
