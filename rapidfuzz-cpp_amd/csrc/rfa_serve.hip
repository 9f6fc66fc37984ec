// rfa_serve.hip -- top-k extraction on the device: the README's extract() loop
// (README.md:153-214: score every choice with a Cached* scorer, keep those
// with score >= score_cutoff) followed by ordering by (score descending,
// choice index ascending) and keeping the first k.
//
// k_serve scores a batch of queries (<= 256 code units, 1-byte units) against
// a resident choice index:
//   * choices are sorted by length and cut into groups of 64 (one wave); a
//     group's characters are stored lane-interleaved, dword d of lane l at
//     data[base + d * 64 + l], so one load per wave reads 256 contiguous bytes;
//   * the PM rows of QPB queries sit in LDS, indexed directly by byte value,
//     word-major ([query][word][byte]): a wave's gather for one word reads
//     64 8-byte entries spread over 256 consecutive ones, as with one word
//     (a [byte][word] layout spaces them 8*WQ bytes apart, so the 64 lanes'
//     addresses fall on 64/WQ bank pairs and conflict WQ times as often);
//   * each wave walks its choices' characters once and advances the
//     recurrence of every query of its workgroup per character (choice
//     outer, queries inner), so choice bytes leave HBM once per batch;
//   * a query only visits the groups whose lengths can reach the cutoff
//     (host-computed group range; pairs outside it score below the cutoff,
//     as in the reference's own length filters LCSseq_impl.hpp:387-389 and
//     Levenshtein_impl.hpp:869-870);
//   * results >= cutoff are appended to the query's candidate list with one
//     atomic per wave; in dense mode (cdist of a few queries against a set,
//     one pair per lane) every result is written to its row instead.
//
// k_select orders candidates: one workgroup per query, thread-local sorted
// lists of KT entries, then block-wide argmax rounds.  A pass ends early when
// the winning thread ran out of listed entries but had dropped some; the next
// pass rescans for entries strictly after the last pick, so any k is exact.
#include "rfa_device.hpp"
#include "rfa_internal.hpp"
#include "rfa_semantics.hpp"
#include "rfa_steps.hpp"

namespace rfa {

__device__ __forceinline__ bool better(double s1, int64_t i1, double s2, int64_t i2)
{
    return s1 > s2 || (s1 == s2 && i1 < i2);
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per-lane recurrence state of one query: WQ 64-bit words as 32-bit halves.
template <int ALGO, int WQ>
struct QState {
    uint32_t a[WQ][2]; // LCS: S; Levenshtein: VP
    uint32_t b[WQ][2]; // Levenshtein: VN

    __device__ __forceinline__ void init()
    {
#pragma unroll
        for (int w = 0; w < WQ; ++w) {
            a[w][0] = a[w][1] = ~0u;
            b[w][0] = b[w][1] = 0u;
        }
    }

    // One text character; row points at the query's PM word 0 for that
    // character, word w at row[w * 256].
    __device__ __forceinline__ void step(const uint64_t* row)
    {
        if constexpr (WQ == 1) {
            const uint64_t X = row[0];
            if constexpr (ALGO == ALGO_LCS)
                lcs_step(X, a[0][0], a[0][1]);
            else
                lev_step(X, a[0][0], a[0][1], b[0][0], b[0][1]);
        }
        else if constexpr (ALGO == ALGO_LCS) {
            // LCSseq_impl.hpp:219-252: S = (S + u + carry) | (S - u), u = S & X
            unsigned long long carry = 0;
#pragma unroll
            for (int w = 0; w < WQ; ++w) {
                const uint64_t X = row[w * 256];
                const uint64_t S = ((uint64_t)a[w][1] << 32) | a[w][0];
                const uint64_t u = S & X;
                unsigned long long co;
                const uint64_t sum = __builtin_addcll(S, u, carry, &co);
                carry = co;
                const uint64_t r = sum | (S & ~X);
                a[w][0] = (uint32_t)r;
                a[w][1] = (uint32_t)(r >> 32);
            }
        }
        else {
            // Levenshtein_impl.hpp:700-736 (block words, carries HP/HN between words)
            uint64_t hp_c = 1, hn_c = 0;
#pragma unroll
            for (int w = 0; w < WQ; ++w) {
                const uint64_t VP = ((uint64_t)a[w][1] << 32) | a[w][0];
                const uint64_t VN = ((uint64_t)b[w][1] << 32) | b[w][0];
                const uint64_t X = row[w * 256] | hn_c;
                const uint64_t D0 = (((X & VP) + VP) ^ VP) | X | VN;
                uint64_t HP = VN | ~(D0 | VP);
                uint64_t HN = D0 & VP;
                const uint64_t hpo = HP >> 63, hno = HN >> 63;
                HP = (HP << 1) | hp_c;
                HN = (HN << 1) | hn_c;
                hp_c = hpo;
                hn_c = hno;
                const uint64_t nVP = HN | ~(D0 | HP);
                const uint64_t nVN = HP & D0;
                a[w][0] = (uint32_t)nVP;
                a[w][1] = (uint32_t)(nVP >> 32);
                b[w][0] = (uint32_t)nVN;
                b[w][1] = (uint32_t)(nVN >> 32);
            }
        }
    }

    // Exact raw value after n text characters against a query of m units:
    // LCS length, or D[m][n] = n + popcount(VP & M) - popcount(VN & M).
    __device__ __forceinline__ uint64_t raw(uint32_t m, uint32_t n) const
    {
        int64_t acc = ALGO == ALGO_LCS ? 0 : (int64_t)n;
#pragma unroll
        for (int w = 0; w < WQ; ++w) {
            const uint32_t lo_bits = (uint32_t)w * 64;
            uint64_t M;
            if (m >= lo_bits + 64)
                M = ~0ull;
            else if (m > lo_bits)
                M = (1ull << (m - lo_bits)) - 1;
            else
                M = 0;
            const uint64_t A = ((uint64_t)a[w][1] << 32) | a[w][0];
            if constexpr (ALGO == ALGO_LCS) {
                acc += __popcll(~A & M);
            }
            else {
                const uint64_t B = ((uint64_t)b[w][1] << 32) | b[w][0];
                acc += (int64_t)__popcll(A & M) - (int64_t)__popcll(B & M);
            }
        }
        return (uint64_t)acc;
    }
};

// DENSE: the cdist-of-a-few-queries mode (every result written); a separate
// instantiation, so the top-k kernels keep their register footprint.
template <int ALGO, int WQ, int QPB, bool DENSE>
__global__ __launch_bounds__(256) void k_serve(ServeArgs a)
{
    extern __shared__ uint64_t pm[]; // [QPB][WQ][256]
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t sub = blockIdx.y;
    const uint32_t q0 = a.q_begin + sub * QPB;
    const uint32_t nqs = min(a.q_end - q0, (uint32_t)QPB);

    for (uint32_t i = tid; i < QPB * 256 * WQ; i += 256) pm[i] = 0;
    __syncthreads();
    for (uint32_t q = 0; q < nqs; ++q) {
        const uint64_t o = a.qoff[q0 + q];
        const uint32_t m = (uint32_t)(a.qoff[q0 + q + 1] - o);
        for (uint32_t i = tid; i < m; i += 256)
            atomicOr((unsigned long long*)&pm[((size_t)q * WQ + (i >> 6)) * 256 + a.qchars[o + i]], 1ull << (i & 63));
    }
    __syncthreads();

    uint32_t qm[QPB], qlo[QPB], qhi[QPB];
#pragma unroll
    for (int q = 0; q < QPB; ++q) {
        if ((uint32_t)q < nqs) {
            qm[q] = (uint32_t)(a.qoff[q0 + q + 1] - a.qoff[q0 + q]);
            qlo[q] = a.qglo[q0 + q];
            qhi[q] = a.qghi[q0 + q];
        }
        else {
            qm[q] = 0;
            qlo[q] = 1;
            qhi[q] = 0;
        }
    }
    const uint32_t glo = a.sub_glo[sub], ghi = a.sub_ghi[sub];

    for (uint32_t g = glo + blockIdx.x * 4 + wave; g < ghi; g += gridDim.x * 4) {
        uint32_t act = 0;
#pragma unroll
        for (int q = 0; q < QPB; ++q)
            if (g >= qlo[q] && g < qhi[q]) act |= 1u << q;
        act = uniform(act);
        if (!act) continue;
        const uint32_t D = a.gdwords[g];
        const uint32_t nmin = a.glen_min[g];
        const uint32_t* gd = a.gdata + a.gbase[g] + lane;
        const uint32_t n = a.slot_len[(size_t)g * 64 + lane];
        const uint32_t cidx = a.slot_idx[(size_t)g * 64 + lane];

        QState<ALGO, WQ> st[QPB];
#pragma unroll
        for (int q = 0; q < QPB; ++q) st[q].init();

        // dwords whose 4 characters exist in every lane: no per-lane masking
        const uint32_t dfull = nmin >> 2;
        uint32_t v = dfull ? gd[0] : 0u;
        for (uint32_t d = 0; d < dfull; ++d) {
            const uint32_t cur = v;
            if (d + 1 < D) v = gd[(size_t)(d + 1) * 64];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint64_t* row = pm + ((cur >> (8 * b)) & 0xffu);
#pragma unroll
                for (int q = 0; q < QPB; ++q)
                    if (act & (1u << q)) st[q].step(row + (size_t)q * 256 * WQ);
            }
        }
        // ragged tail: characters past a lane's length leave its state alone
        for (uint32_t d = dfull; d < D; ++d) {
            const uint32_t cur = gd[(size_t)d * 64];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t j = d * 4 + b;
                if (j < n) {
                    const uint64_t* row = pm + ((cur >> (8 * b)) & 0xffu);
#pragma unroll
                    for (int q = 0; q < QPB; ++q)
                        if (act & (1u << q)) st[q].step(row + (size_t)q * 256 * WQ);
                }
            }
        }

#pragma unroll
        for (int q = 0; q < QPB; ++q) {
            if (!(act & (1u << q))) continue;
            RawResult r{st[q].raw(qm[q], n), 0, 0};
            if constexpr (DENSE) {
                const uint32_t qr = a.qrow[q0 + q];
                ScoreSpec sp = a.spec;
                if (a.qcut) sp.cutoff_u = a.qcut[qr];
                if (cidx != 0xffffffffu) a.dense_out[(uint64_t)qr * a.dense_stride + cidx] = finalize_bits(sp, qm[q], n, r);
                continue;
            }
            const double score = __builtin_bit_cast(double, finalize_bits(a.spec, qm[q], n, r));
            const bool pass = cidx != 0xffffffffu && score >= a.cutoff;
            const uint64_t bal = ballot(pass);
            if (!bal) continue;
            const uint32_t leader = (uint32_t)__builtin_ctzll(bal);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&a.counts[q0 + q], (uint32_t)__popcll(bal));
            base = readlane(base, leader);
            if (pass) {
                const uint32_t pos = base + mbcnt64(bal);
                if (pos < a.cap) {
                    a.cand_score[(size_t)(q0 + q) * a.cap + pos] = score;
                    a.cand_index[(size_t)(q0 + q) * a.cap + pos] = (int64_t)cidx;
                }
            }
        }
    }
}

// ---- selection ----------------------------------------------------------------

constexpr int KT = 8;

struct LocalList {
    double s[KT];
    int64_t i[KT];
    uint32_t cnt;
    bool dropped;

    __device__ __forceinline__ void clear()
    {
        cnt = 0;
        dropped = false;
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            s[t] = -INFINITY;
            i[t] = INT64_MAX;
        }
    }
    __device__ __forceinline__ void insert(double ns, int64_t ni)
    {
        if (cnt == KT && !better(ns, ni, s[KT - 1], i[KT - 1])) {
            dropped = true;
            return;
        }
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            if ((uint32_t)t < cnt) {
                if (better(ns, ni, s[t], i[t])) {
                    const double ts = s[t];
                    const int64_t ti = i[t];
                    s[t] = ns;
                    i[t] = ni;
                    ns = ts;
                    ni = ti;
                }
            }
            else if ((uint32_t)t == cnt) {
                s[t] = ns;
                i[t] = ni;
            }
        }
        if (cnt < KT)
            ++cnt;
        else
            dropped = true;
    }
    __device__ __forceinline__ void pop()
    {
#pragma unroll
        for (int t = 0; t + 1 < KT; ++t) {
            s[t] = s[t + 1];
            i[t] = i[t + 1];
        }
        s[KT - 1] = -INFINITY;
        i[KT - 1] = INT64_MAX;
        --cnt;
    }
};

__global__ __launch_bounds__(256) void k_select(SelectArgs a)
{
    __shared__ double ws[4];
    __shared__ int64_t wi[4];
    __shared__ uint32_t wo[4];
    __shared__ double bs_sh;
    __shared__ int64_t bi_sh;
    __shared__ uint32_t bo_sh;
    __shared__ uint32_t restart_sh;
    const uint32_t q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t lcount = a.lcount ? min(a.lcount[q], a.lcap) : a.lfixed;
    const double* ls = a.lscore ? a.lscore + (size_t)q * a.lstride : nullptr;
    const int64_t* li = a.lindex ? a.lindex + (size_t)q * a.lstride : nullptr;
    const double* ds = a.dscore ? a.dscore + (size_t)q * a.dstride : nullptr;
    const uint64_t dn = a.dscore ? a.dn : 0;
    double* os = a.out_score + (size_t)q * a.k;
    int64_t* oi = a.out_index + (size_t)q * a.k;

    uint32_t produced = 0;
    bool have_last = false;
    double last_s = 0;
    int64_t last_i = 0;
    LocalList L;
    while (produced < a.k) {
        L.clear();
        for (uint32_t t = tid; t < lcount; t += 256) {
            const double s = ls[t];
            const int64_t i = li[t];
            if (i < 0 || !(s >= a.cutoff)) continue;
            if (have_last && !better(last_s, last_i, s, i)) continue;
            L.insert(s, i);
        }
        for (uint64_t t = tid; t < dn; t += 256) {
            const double s = ds[t];
            const int64_t i = a.dbase + (int64_t)t;
            if (!(s >= a.cutoff)) continue;
            if (have_last && !better(last_s, last_i, s, i)) continue;
            L.insert(s, i);
        }
        bool restart = false;
        while (produced < a.k && !restart) {
            double s = L.cnt ? L.s[0] : -INFINITY;
            int64_t i = L.cnt ? L.i[0] : INT64_MAX;
            uint32_t o = tid;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double s2 = __shfl_xor(s, off, 64);
                const int64_t i2 = __shfl_xor(i, off, 64);
                const uint32_t o2 = __shfl_xor(o, off, 64);
                if (better(s2, i2, s, i)) {
                    s = s2;
                    i = i2;
                    o = o2;
                }
            }
            if (lane == 0) {
                ws[wave] = s;
                wi[wave] = i;
                wo[wave] = o;
            }
            __syncthreads();
            if (tid == 0) {
                double bs = ws[0];
                int64_t bi = wi[0];
                uint32_t bo = wo[0];
                for (int w = 1; w < 4; ++w) {
                    if (better(ws[w], wi[w], bs, bi)) {
                        bs = ws[w];
                        bi = wi[w];
                        bo = wo[w];
                    }
                }
                bs_sh = bs;
                bi_sh = bi;
                bo_sh = bo;
                restart_sh = 0;
            }
            __syncthreads();
            const double bs = bs_sh;
            const int64_t bi = bi_sh;
            const uint32_t bo = bo_sh;
            if (bi == INT64_MAX) { // nothing left
                for (uint32_t r = produced + tid; r < a.k; r += 256) {
                    os[r] = 0.0;
                    oi[r] = -1;
                }
                return;
            }
            if (tid == 0) {
                os[produced] = bs;
                oi[produced] = bi;
            }
            ++produced;
            have_last = true;
            last_s = bs;
            last_i = bi;
            if (tid == bo) {
                L.pop();
                if (L.cnt == 0 && L.dropped) restart_sh = 1;
            }
            __syncthreads();
            restart = restart_sh != 0;
            __syncthreads();
        }
    }
}

// ---- launchers ------------------------------------------------------------------

template <int ALGO, int WQ, int QPB, bool DENSE>
static hipError_t launch_serve_t(const ServeArgs& a, uint32_t gx, uint32_t nsub, hipStream_t s)
{
    const size_t lds = (size_t)QPB * 256 * WQ * 8;
    hipLaunchKernelGGL((k_serve<ALGO, WQ, QPB, DENSE>), dim3(gx, nsub), dim3(256), lds, s, a);
    return hipGetLastError();
}

uint32_t serve_queries_per_block(uint32_t wq) { return wq <= 2 ? 8 : 4; }

template <bool DENSE>
static hipError_t launch_serve_d(int algo, uint32_t wq, const ServeArgs& a, uint32_t gx, uint32_t nsub, hipStream_t s)
{
    if (algo == ALGO_LCS) {
        if (wq <= 1) return launch_serve_t<ALGO_LCS, 1, 8, DENSE>(a, gx, nsub, s);
        if (wq <= 2) return launch_serve_t<ALGO_LCS, 2, 8, DENSE>(a, gx, nsub, s);
        return launch_serve_t<ALGO_LCS, 4, 4, DENSE>(a, gx, nsub, s);
    }
    if (wq <= 1) return launch_serve_t<ALGO_LEV, 1, 8, DENSE>(a, gx, nsub, s);
    if (wq <= 2) return launch_serve_t<ALGO_LEV, 2, 8, DENSE>(a, gx, nsub, s);
    return launch_serve_t<ALGO_LEV, 4, 4, DENSE>(a, gx, nsub, s);
}

hipError_t launch_serve(int algo, uint32_t wq, const ServeArgs& a, uint32_t gx, uint32_t nsub, hipStream_t s)
{
    return a.dense_out ? launch_serve_d<true>(algo, wq, a, gx, nsub, s) : launch_serve_d<false>(algo, wq, a, gx, nsub, s);
}

hipError_t launch_select(const SelectArgs& a, uint32_t nq, hipStream_t s)
{
    hipLaunchKernelGGL(k_select, dim3(nq), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace rfa
