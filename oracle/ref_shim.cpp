// ref_shim.cpp -- extern "C" wrapper around the reference's own headers.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (-I <reference root>), output only into
// oracle/_ref/.  It gives the tests a second, independent checker (the real
// reference library) and gives bench.py a "reference" CPU baseline.  The
// product library never links or loads it.
//
// Entry points mirror the reference call shapes:
//   Cached<Metric><CharT>(s1).<kind>(s2, cutoff)         (e.g. Levenshtein.hpp:421-482)
//   experimental::Multi<Metric><MaxLen>(n).insert / .<kind>(scores, n, s2, cutoff)
//                                                        (e.g. Levenshtein.hpp:297-419)
#include <rapidfuzz/distance.hpp>
#include <rapidfuzz/fuzz.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

enum { LEV = 0, INDEL = 1, LCS = 2, OSA = 3, JARO = 4, JW = 5, RATIO = 6, HAMMING = 7, PREFIX = 8, POSTFIX = 9, PARTIAL = 10, TOKEN_SORT = 11, PARTIAL_TOKEN_SORT = 12, QRATIO = 13, TOKEN_SET = 14, TOKEN_RATIO = 15, PTOKEN_SET = 16, PTOKEN_RATIO = 17, DAMERAU = 18, WRATIO = 19 };
enum { DIST = 0, SIM = 1, NDIST = 2, NSIM = 3 };

thread_local std::string g_err;

template <typename CharT>
std::vector<CharT> take(const uint32_t* s, size_t n)
{
    std::vector<CharT> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = static_cast<CharT>(s[i]);
    return v;
}

template <typename Scorer, typename It>
int run_int_metric(const Scorer& sc, int kind, It b, It e, uint64_t cu, double cd, uint64_t* ou, double* od)
{
    switch (kind) {
    case DIST: *ou = sc.distance(b, e, static_cast<size_t>(cu)); return 0;
    case SIM: *ou = sc.similarity(b, e, static_cast<size_t>(cu)); return 0;
    case NDIST: *od = sc.normalized_distance(b, e, cd); return 1;
    case NSIM: *od = sc.normalized_similarity(b, e, cd); return 1;
    }
    return -1;
}

template <typename Scorer, typename It>
int run_fp_metric(const Scorer& sc, int kind, It b, It e, double cd, double* od)
{
    switch (kind) {
    case DIST: *od = sc.distance(b, e, cd); return 1;
    case SIM: *od = sc.similarity(b, e, cd); return 1;
    case NDIST: *od = sc.normalized_distance(b, e, cd); return 1;
    case NSIM: *od = sc.normalized_similarity(b, e, cd); return 1;
    }
    return -1;
}

// A cached scorer for one query, dispatching on the metric at run time.
template <typename CharT>
struct AnyCached {
    int metric;
    std::vector<CharT> s1;
    rapidfuzz::LevenshteinWeightTable w;
    double pw;
    rapidfuzz::CachedLevenshtein<CharT>* lev = nullptr;
    rapidfuzz::CachedIndel<CharT>* indel = nullptr;
    rapidfuzz::CachedLCSseq<CharT>* lcs = nullptr;
    rapidfuzz::CachedOSA<CharT>* osa = nullptr;
    rapidfuzz::CachedJaro<CharT>* jaro = nullptr;
    rapidfuzz::CachedJaroWinkler<CharT>* jw = nullptr;
    rapidfuzz::fuzz::CachedRatio<CharT>* ratio = nullptr;
    rapidfuzz::CachedHamming<CharT>* ham = nullptr;
    rapidfuzz::CachedPrefix<CharT>* pre = nullptr;
    rapidfuzz::CachedPostfix<CharT>* post = nullptr;
    rapidfuzz::fuzz::CachedPartialRatio<CharT>* partial = nullptr;
    rapidfuzz::fuzz::CachedTokenSortRatio<CharT>* tsort = nullptr;
    rapidfuzz::fuzz::CachedPartialTokenSortRatio<CharT>* ptsort = nullptr;
    rapidfuzz::fuzz::CachedQRatio<CharT>* qratio = nullptr;
    rapidfuzz::fuzz::CachedTokenSetRatio<CharT>* tset = nullptr;
    rapidfuzz::fuzz::CachedTokenRatio<CharT>* tratio = nullptr;
    rapidfuzz::fuzz::CachedPartialTokenSetRatio<CharT>* ptset = nullptr;
    rapidfuzz::fuzz::CachedPartialTokenRatio<CharT>* ptratio = nullptr;
    rapidfuzz::experimental::CachedDamerauLevenshtein<CharT>* dl = nullptr;
    rapidfuzz::fuzz::CachedWRatio<CharT>* wratio = nullptr;

    AnyCached(int m, std::vector<CharT> q, rapidfuzz::LevenshteinWeightTable w_, double pw_)
        : metric(m), s1(std::move(q)), w(w_), pw(pw_)
    {
        switch (metric) {
        case LEV: lev = new rapidfuzz::CachedLevenshtein<CharT>(s1, w); break;
        case INDEL: indel = new rapidfuzz::CachedIndel<CharT>(s1); break;
        case LCS: lcs = new rapidfuzz::CachedLCSseq<CharT>(s1); break;
        case OSA: osa = new rapidfuzz::CachedOSA<CharT>(s1); break;
        case JARO: jaro = new rapidfuzz::CachedJaro<CharT>(s1); break;
        case JW: jw = new rapidfuzz::CachedJaroWinkler<CharT>(s1, pw); break;
        case RATIO: ratio = new rapidfuzz::fuzz::CachedRatio<CharT>(s1); break;
        case HAMMING: ham = new rapidfuzz::CachedHamming<CharT>(s1, true); break;
        case PREFIX: pre = new rapidfuzz::CachedPrefix<CharT>(s1); break;
        case POSTFIX: post = new rapidfuzz::CachedPostfix<CharT>(s1); break;
        case PARTIAL: partial = new rapidfuzz::fuzz::CachedPartialRatio<CharT>(s1); break;
        case TOKEN_SORT: tsort = new rapidfuzz::fuzz::CachedTokenSortRatio<CharT>(s1); break;
        case PARTIAL_TOKEN_SORT: ptsort = new rapidfuzz::fuzz::CachedPartialTokenSortRatio<CharT>(s1); break;
        case QRATIO: qratio = new rapidfuzz::fuzz::CachedQRatio<CharT>(s1); break;
        case TOKEN_SET: tset = new rapidfuzz::fuzz::CachedTokenSetRatio<CharT>(s1); break;
        case TOKEN_RATIO: tratio = new rapidfuzz::fuzz::CachedTokenRatio<CharT>(s1); break;
        case PTOKEN_SET: ptset = new rapidfuzz::fuzz::CachedPartialTokenSetRatio<CharT>(s1); break;
        case PTOKEN_RATIO: ptratio = new rapidfuzz::fuzz::CachedPartialTokenRatio<CharT>(s1); break;
        case DAMERAU: dl = new rapidfuzz::experimental::CachedDamerauLevenshtein<CharT>(s1); break;
        case WRATIO: wratio = new rapidfuzz::fuzz::CachedWRatio<CharT>(s1); break;
        default: throw std::invalid_argument("unknown metric");
        }
    }
    ~AnyCached()
    {
        delete ham;
        delete pre;
        delete post;
        delete partial;
        delete tsort;
        delete ptsort;
        delete qratio;
        delete tset;
        delete tratio;
        delete ptset;
        delete ptratio;
        delete dl;
        delete wratio;
        delete lev;
        delete indel;
        delete lcs;
        delete osa;
        delete jaro;
        delete jw;
        delete ratio;
    }
    template <typename It>
    int score(int kind, It b, It e, uint64_t cu, double cd, uint64_t* ou, double* od) const
    {
        switch (metric) {
        case LEV: return run_int_metric(*lev, kind, b, e, cu, cd, ou, od);
        case INDEL: return run_int_metric(*indel, kind, b, e, cu, cd, ou, od);
        case LCS: return run_int_metric(*lcs, kind, b, e, cu, cd, ou, od);
        case OSA: return run_int_metric(*osa, kind, b, e, cu, cd, ou, od);
        case HAMMING: return run_int_metric(*ham, kind, b, e, cu, cd, ou, od);
        case PREFIX: return run_int_metric(*pre, kind, b, e, cu, cd, ou, od);
        case POSTFIX: return run_int_metric(*post, kind, b, e, cu, cd, ou, od);
        case DAMERAU: return run_int_metric(*dl, kind, b, e, cu, cd, ou, od);
        case JARO: return run_fp_metric(*jaro, kind, b, e, cd, od);
        case JW: return run_fp_metric(*jw, kind, b, e, cd, od);
        case RATIO:
            if (kind != SIM) return -1;
            *od = ratio->similarity(b, e, cd);
            return 1;
        case PARTIAL:
            if (kind != SIM) return -1;
            *od = partial->similarity(b, e, cd);
            return 1;
        case TOKEN_SORT:
            if (kind != SIM) return -1;
            *od = tsort->similarity(b, e, cd);
            return 1;
        case PARTIAL_TOKEN_SORT:
            if (kind != SIM) return -1;
            *od = ptsort->similarity(b, e, cd);
            return 1;
        case QRATIO:
            if (kind != SIM) return -1;
            *od = qratio->similarity(b, e, cd);
            return 1;
        case TOKEN_SET:
            if (kind != SIM) return -1;
            *od = tset->similarity(b, e, cd);
            return 1;
        case TOKEN_RATIO:
            if (kind != SIM) return -1;
            *od = tratio->similarity(b, e, cd);
            return 1;
        case PTOKEN_SET:
            if (kind != SIM) return -1;
            *od = ptset->similarity(b, e, cd);
            return 1;
        case PTOKEN_RATIO:
            if (kind != SIM) return -1;
            *od = ptratio->similarity(b, e, cd);
            return 1;
        case WRATIO:
            if (kind != SIM) return -1;
            *od = wratio->similarity(b, e, cd);
            return 1;
        }
        return -1;
    }
};

template <typename CharT>
int cdist_cached_impl(int metric, int kind, const uint32_t* q, const uint64_t* qoff, size_t nq,
                      const uint32_t* c, const uint64_t* coff, size_t nc, uint64_t ins, uint64_t del,
                      uint64_t rep, double pw, uint64_t cu, double cd, uint64_t* ou, double* od)
{
    std::vector<std::vector<CharT>> choices(nc);
    for (size_t j = 0; j < nc; ++j) choices[j] = take<CharT>(c + coff[j], coff[j + 1] - coff[j]);
    int r = -1;
    for (size_t i = 0; i < nq; ++i) {
        AnyCached<CharT> sc(metric, take<CharT>(q + qoff[i], qoff[i + 1] - qoff[i]),
                            {static_cast<size_t>(ins), static_cast<size_t>(del), static_cast<size_t>(rep)},
                            pw);
        for (size_t j = 0; j < nc; ++j) {
            uint64_t u = 0;
            double d = 0;
            r = sc.score(kind, choices[j].data(), choices[j].data() + choices[j].size(), cu, cd, &u, &d);
            if (r < 0) return r;
            if (r == 0)
                ou[i * nc + j] = u;
            else
                od[i * nc + j] = d;
        }
    }
    return r;
}

// ---- experimental::Multi* ----
template <int MaxLen>
struct AnyMulti {
    int metric;
    size_t n;
    rapidfuzz::experimental::MultiLevenshtein<MaxLen>* lev = nullptr;
    rapidfuzz::experimental::MultiIndel<MaxLen>* indel = nullptr;
    rapidfuzz::experimental::MultiLCSseq<MaxLen>* lcs = nullptr;
    rapidfuzz::experimental::MultiOSA<MaxLen>* osa = nullptr;
    rapidfuzz::experimental::MultiJaro<MaxLen>* jaro = nullptr;
    rapidfuzz::experimental::MultiJaroWinkler<MaxLen>* jw = nullptr;
    rapidfuzz::fuzz::experimental::MultiRatio<MaxLen>* ratio = nullptr;

    AnyMulti(int m, size_t n_, double pw, rapidfuzz::LevenshteinWeightTable w = {1, 1, 1}) : metric(m), n(n_)
    {
        switch (metric) {
        case LEV: lev = new rapidfuzz::experimental::MultiLevenshtein<MaxLen>(n, w); break;
        case INDEL: indel = new rapidfuzz::experimental::MultiIndel<MaxLen>(n); break;
        case LCS: lcs = new rapidfuzz::experimental::MultiLCSseq<MaxLen>(n); break;
        case OSA: osa = new rapidfuzz::experimental::MultiOSA<MaxLen>(n); break;
        case JARO: jaro = new rapidfuzz::experimental::MultiJaro<MaxLen>(n); break;
        case JW: jw = new rapidfuzz::experimental::MultiJaroWinkler<MaxLen>(n, pw); break;
        case RATIO: ratio = new rapidfuzz::fuzz::experimental::MultiRatio<MaxLen>(n); break;
        default: throw std::invalid_argument("unknown metric");
        }
    }
    ~AnyMulti()
    {
        delete lev;
        delete indel;
        delete lcs;
        delete osa;
        delete jaro;
        delete jw;
        delete ratio;
    }
    size_t result_count() const
    {
        switch (metric) {
        case LEV: return lev->result_count();
        case INDEL: return indel->result_count();
        case LCS: return lcs->result_count();
        case OSA: return osa->result_count();
        case JARO: return jaro->result_count();
        case JW: return jw->result_count();
        case RATIO: return ratio->result_count();
        }
        return 0;
    }
    template <typename S>
    void insert(const S& s)
    {
        switch (metric) {
        case LEV: lev->insert(s); break;
        case INDEL: indel->insert(s); break;
        case LCS: lcs->insert(s); break;
        case OSA: osa->insert(s); break;
        case JARO: jaro->insert(s); break;
        case JW: jw->insert(s); break;
        case RATIO: ratio->insert(s); break;
        }
    }
    template <typename M, typename It>
    static int run_int(const M& m, int kind, size_t* su, double* sd, size_t cnt, It b, It e, uint64_t cu,
                       double cd)
    {
        switch (kind) {
        case DIST: m.distance(su, cnt, b, e, static_cast<size_t>(cu)); return 0;
        case SIM: m.similarity(su, cnt, b, e, static_cast<size_t>(cu)); return 0;
        case NDIST: m.normalized_distance(sd, cnt, b, e, cd); return 1;
        case NSIM: m.normalized_similarity(sd, cnt, b, e, cd); return 1;
        }
        return -1;
    }
    template <typename M, typename It>
    static int run_fp(const M& m, int kind, double* sd, size_t cnt, It b, It e, double cd)
    {
        switch (kind) {
        case DIST: m.distance(sd, cnt, b, e, cd); return 1;
        case SIM: m.similarity(sd, cnt, b, e, cd); return 1;
        case NDIST: m.normalized_distance(sd, cnt, b, e, cd); return 1;
        case NSIM: m.normalized_similarity(sd, cnt, b, e, cd); return 1;
        }
        return -1;
    }
    template <typename It>
    int score(int kind, size_t* su, double* sd, size_t cnt, It b, It e, uint64_t cu, double cd) const
    {
        switch (metric) {
        case LEV: return run_int(*lev, kind, su, sd, cnt, b, e, cu, cd);
        case INDEL: return run_int(*indel, kind, su, sd, cnt, b, e, cu, cd);
        case LCS: return run_int(*lcs, kind, su, sd, cnt, b, e, cu, cd);
        case OSA: return run_int(*osa, kind, su, sd, cnt, b, e, cu, cd);
        case JARO: return run_fp(*jaro, kind, sd, cnt, b, e, cd);
        case JW: return run_fp(*jw, kind, sd, cnt, b, e, cd);
        case RATIO:
            if (kind != SIM) return -1;
            ratio->similarity(sd, cnt, b, e, cd);
            return 1;
        }
        return -1;
    }
};

template <int MaxLen, typename CharT>
int multi_cdist_impl(int metric, int kind, const uint32_t* q, const uint64_t* qoff, size_t nq,
                     const uint32_t* c, const uint64_t* coff, size_t nc, double pw, uint64_t cu,
                     double cd, uint64_t* ou, double* od, rapidfuzz::LevenshteinWeightTable w = {1, 1, 1})
{
    AnyMulti<MaxLen> m(metric, nq, pw, w);
    for (size_t i = 0; i < nq; ++i) m.insert(take<CharT>(q + qoff[i], qoff[i + 1] - qoff[i]));
    size_t rc = m.result_count();
    std::vector<size_t> su(rc);
    std::vector<double> sd(rc);
    int r = -1;
    for (size_t j = 0; j < nc; ++j) {
        auto s2 = take<CharT>(c + coff[j], coff[j + 1] - coff[j]);
        r = m.score(kind, su.data(), sd.data(), rc, s2.data(), s2.data() + s2.size(), cu, cd);
        if (r < 0) return r;
        for (size_t i = 0; i < nq; ++i) {
            if (r == 0)
                ou[i * nc + j] = su[i];
            else
                od[i * nc + j] = sd[i];
        }
    }
    return r;
}

// ---- threaded u8 cdist used as the CPU baseline ----
struct U8Set {
    const uint8_t* chars;
    const uint64_t* off;
    size_t n;
};

int cdist_cached_u8_threaded(int metric, int kind, U8Set q, U8Set c, rapidfuzz::LevenshteinWeightTable w, double pw,
                             uint64_t cu, double cd, uint64_t* ou, double* od, int nthreads)
{
    std::atomic<size_t> next(0);
    std::atomic<int> status(0);
    auto worker = [&]() {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= q.n) break;
            std::vector<uint8_t> s1(q.chars + q.off[i], q.chars + q.off[i + 1]);
            AnyCached<uint8_t> sc(metric, s1, w, pw);
            for (size_t j = 0; j < c.n; ++j) {
                const uint8_t* b = c.chars + c.off[j];
                const uint8_t* e = c.chars + c.off[j + 1];
                uint64_t u = 0;
                double d = 0;
                int r = sc.score(kind, b, e, cu, cd, &u, &d);
                if (r < 0) {
                    status = -1;
                    return;
                }
                if (r == 0)
                    ou[i * c.n + j] = u;
                else
                    od[i * c.n + j] = d;
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
    return status.load();
}

// Multi* over length buckets (MaxLen 8/16/32/64), threads split the choices.
template <int MaxLen>
void multi_bucket_u8(int metric, int kind, const std::vector<size_t>& idx, U8Set q, U8Set c, double pw,
                     uint64_t cu, double cd, uint64_t* ou, double* od, int nthreads)
{
    if (idx.empty()) return;
    AnyMulti<MaxLen> m(metric, idx.size(), pw);
    for (size_t k = 0; k < idx.size(); ++k) {
        size_t i = idx[k];
        m.insert(std::vector<uint8_t>(q.chars + q.off[i], q.chars + q.off[i + 1]));
    }
    size_t rc = m.result_count();
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        std::vector<size_t> su(rc);
        std::vector<double> sd(rc);
        const size_t chunk = 256;
        for (;;) {
            size_t j0 = next.fetch_add(chunk);
            if (j0 >= c.n) break;
            size_t j1 = std::min(c.n, j0 + chunk);
            for (size_t j = j0; j < j1; ++j) {
                int r = m.score(kind, su.data(), sd.data(), rc, c.chars + c.off[j], c.chars + c.off[j + 1], cu,
                                cd);
                for (size_t k = 0; k < idx.size(); ++k) {
                    if (r == 0)
                        ou[idx[k] * c.n + j] = su[k];
                    else
                        od[idx[k] * c.n + j] = sd[k];
                }
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
}

} // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Cached* cdist over arbitrary code units; char_width picks CharT (1, 2 or 4; -1 for char).
int ref_cdist_cached(int metric, int kind, int char_width, const uint32_t* q, const uint64_t* qoff,
                     size_t nq, const uint32_t* c, const uint64_t* coff, size_t nc, uint64_t ins,
                     uint64_t del, uint64_t rep, double pw, uint64_t cu, double cd, uint64_t* ou,
                     double* od)
{
    try {
        if (char_width == -1) // `char`: signed 8-bit units
            return cdist_cached_impl<char>(metric, kind, q, qoff, nq, c, coff, nc, ins, del, rep, pw, cu, cd, ou,
                                           od);
        if (char_width == 1)
            return cdist_cached_impl<uint8_t>(metric, kind, q, qoff, nq, c, coff, nc, ins, del, rep, pw, cu,
                                              cd, ou, od);
        if (char_width == 2)
            return cdist_cached_impl<uint16_t>(metric, kind, q, qoff, nq, c, coff, nc, ins, del, rep, pw,
                                               cu, cd, ou, od);
        return cdist_cached_impl<uint32_t>(metric, kind, q, qoff, nq, c, coff, nc, ins, del, rep, pw, cu,
                                           cd, ou, od);
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// experimental::Multi<Metric><max_len> cdist: all queries inserted into one
// scorer; ins/del/rep are MultiLevenshtein's weights (other metrics ignore them).
int ref_cdist_multi_w(int metric, int kind, int max_len, const uint32_t* q, const uint64_t* qoff, size_t nq,
                      const uint32_t* c, const uint64_t* coff, size_t nc, uint64_t ins, uint64_t del, uint64_t rep,
                      double pw, uint64_t cu, double cd, uint64_t* ou, double* od)
{
    const rapidfuzz::LevenshteinWeightTable w{static_cast<size_t>(ins), static_cast<size_t>(del),
                                              static_cast<size_t>(rep)};
    try {
        switch (max_len) {
        case 8: return multi_cdist_impl<8, uint32_t>(metric, kind, q, qoff, nq, c, coff, nc, pw, cu, cd, ou, od, w);
        case 16:
            return multi_cdist_impl<16, uint32_t>(metric, kind, q, qoff, nq, c, coff, nc, pw, cu, cd, ou, od, w);
        case 32:
            return multi_cdist_impl<32, uint32_t>(metric, kind, q, qoff, nq, c, coff, nc, pw, cu, cd, ou, od, w);
        case 64:
            return multi_cdist_impl<64, uint32_t>(metric, kind, q, qoff, nq, c, coff, nc, pw, cu, cd, ou, od, w);
        }
        g_err = "max_len must be 8, 16, 32 or 64";
        return -1;
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

int ref_cdist_multi(int metric, int kind, int max_len, const uint32_t* q, const uint64_t* qoff, size_t nq,
                    const uint32_t* c, const uint64_t* coff, size_t nc, double pw, uint64_t cu, double cd,
                    uint64_t* ou, double* od)
{
    return ref_cdist_multi_w(metric, kind, max_len, q, qoff, nq, c, coff, nc, 1, 1, 1, pw, cu, cd, ou, od);
}

// CPU baseline: Cached* per query, queries spread over nthreads std::threads.
int ref_cdist_cached_u8_mt(int metric, int kind, const uint8_t* q, const uint64_t* qoff, size_t nq,
                           const uint8_t* c, const uint64_t* coff, size_t nc, double pw, uint64_t cu,
                           double cd, uint64_t* ou, double* od, int nthreads)
{
    try {
        return cdist_cached_u8_threaded(metric, kind, {q, qoff, nq}, {c, coff, nc}, {1, 1, 1}, pw, cu, cd, ou,
                                        od, nthreads);
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// The same with Levenshtein weights.
int ref_cdist_cached_u8_mt_w(int metric, int kind, const uint8_t* q, const uint64_t* qoff, size_t nq,
                             const uint8_t* c, const uint64_t* coff, size_t nc, uint64_t ins, uint64_t del,
                             uint64_t rep, double pw, uint64_t cu, double cd, uint64_t* ou, double* od,
                             int nthreads)
{
    try {
        return cdist_cached_u8_threaded(metric, kind, {q, qoff, nq}, {c, coff, nc}, {ins, del, rep}, pw, cu, cd,
                                        ou, od, nthreads);
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// CPU baseline: AVX2 Multi* with queries bucketed by MaxLen 8/16/32/64.
int ref_cdist_multi_u8_mt(int metric, int kind, const uint8_t* q, const uint64_t* qoff, size_t nq,
                          const uint8_t* c, const uint64_t* coff, size_t nc, double pw, uint64_t cu,
                          double cd, uint64_t* ou, double* od, int nthreads)
{
    try {
        std::vector<size_t> b8, b16, b32, b64;
        for (size_t i = 0; i < nq; ++i) {
            size_t l = qoff[i + 1] - qoff[i];
            if (l <= 8)
                b8.push_back(i);
            else if (l <= 16)
                b16.push_back(i);
            else if (l <= 32)
                b32.push_back(i);
            else if (l <= 64)
                b64.push_back(i);
            else {
                g_err = "query longer than 64";
                return -1;
            }
        }
        U8Set qs{q, qoff, nq}, cs{c, coff, nc};
        multi_bucket_u8<8>(metric, kind, b8, qs, cs, pw, cu, cd, ou, od, nthreads);
        multi_bucket_u8<16>(metric, kind, b16, qs, cs, pw, cu, cd, ou, od, nthreads);
        multi_bucket_u8<32>(metric, kind, b32, qs, cs, pw, cu, cd, ou, od, nthreads);
        multi_bucket_u8<64>(metric, kind, b64, qs, cs, pw, cu, cd, ou, od, nthreads);
        return 0;
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// CPU baseline for independent pairs: free functions per pair.
int ref_pairs_u32_mt(const int32_t* metrics, const uint32_t* a, const uint64_t* aoff, const uint32_t* b,
                     const uint64_t* boff, size_t n, uint64_t* out, int nthreads)
{
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        const size_t chunk = 64;
        for (;;) {
            size_t i0 = next.fetch_add(chunk);
            if (i0 >= n) break;
            size_t i1 = std::min(n, i0 + chunk);
            for (size_t i = i0; i < i1; ++i) {
                if (metrics[i] == LEV)
                    out[i] = rapidfuzz::levenshtein_distance(a + aoff[i], a + aoff[i + 1], b + boff[i],
                                                             b + boff[i + 1]);
                else
                    out[i] = rapidfuzz::indel_distance(a + aoff[i], a + aoff[i + 1], b + boff[i], b + boff[i + 1]);
            }
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
    return 0;
}

// CPU baseline for 1-vs-N and banded scoring: CachedLevenshtein<uint8_t> per
// query with that query's cutoff (Levenshtein.hpp:421-482), threads over
// (query, chunk of 256 choices) tasks.  out[i * nc + j].
int ref_lev_cutoff_u8_mt(const uint8_t* q, const uint64_t* qoff, size_t nq, const uint8_t* c, const uint64_t* coff,
                         size_t nc, const uint64_t* cutoffs, uint64_t* out, int nthreads)
{
    const size_t chunk = 256, per_q = (nc + chunk - 1) / chunk;
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        for (;;) {
            const size_t t = next.fetch_add(1);
            if (t >= nq * per_q) break;
            const size_t i = t / per_q, j0 = (t % per_q) * chunk, j1 = std::min(nc, j0 + chunk);
            rapidfuzz::CachedLevenshtein<uint8_t> scorer(q + qoff[i], q + qoff[i + 1]);
            for (size_t j = j0; j < j1; ++j)
                out[i * nc + j] = scorer.distance(c + coff[j], c + coff[j + 1], static_cast<size_t>(cutoffs[i]));
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
    return 0;
}

// CPU baseline for serving: the README's extract() loop (README.md:153-214)
// with fuzz::CachedRatio, then the k best by (score desc, index asc).  One
// request per thread at a time; seconds[i] is request i's own compute time.
int ref_extract_ratio_u8_mt(const uint8_t* q, const uint64_t* qoff, size_t nq, const uint8_t* c,
                            const uint64_t* coff, size_t nc, double cutoff, size_t k, int64_t* out_index,
                            double* out_score, double* seconds, int nthreads)
{
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        std::vector<std::pair<double, size_t>> hits;
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= nq) break;
            const auto t0 = std::chrono::steady_clock::now();
            rapidfuzz::fuzz::CachedRatio<uint8_t> scorer(q + qoff[i], q + qoff[i + 1]);
            hits.clear();
            for (size_t j = 0; j < nc; ++j) {
                const double s = scorer.similarity(c + coff[j], c + coff[j + 1], cutoff);
                if (s >= cutoff) hits.emplace_back(s, j);
            }
            const size_t kk = std::min(k, hits.size());
            std::partial_sort(hits.begin(), hits.begin() + kk, hits.end(), [](const auto& a, const auto& b) {
                return a.first > b.first || (a.first == b.first && a.second < b.second);
            });
            for (size_t r = 0; r < k; ++r) {
                out_index[i * k + r] = r < kk ? (int64_t)hits[r].second : -1;
                out_score[i * k + r] = r < kk ? hits[r].first : 0.0;
            }
            seconds[i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    };
    std::vector<std::thread> ts;
    for (int t = 0; t < nthreads; ++t) ts.emplace_back(worker);
    for (auto& t : ts) t.join();
    return 0;
}

// levenshtein_editops (metric 0) or lcs_seq_editops (otherwise) of one pair of 32-bit
// strings; returns the count, writes at most l1 + l2 operations.
int64_t ref_editops(int metric, const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, int32_t* type,
                    uint64_t* src_pos, uint64_t* dest_pos)
{
    try {
        std::vector<uint32_t> a(s1, s1 + l1), b(s2, s2 + l2);
        rapidfuzz::Editops ops = metric == 0   ? rapidfuzz::levenshtein_editops(a, b)
                                 : metric == 7 ? rapidfuzz::hamming_editops(a, b)
                                               : rapidfuzz::lcs_seq_editops(a, b);
        for (size_t i = 0; i < ops.size(); ++i) {
            type[i] = static_cast<int32_t>(ops[i].type);
            src_pos[i] = ops[i].src_pos;
            dest_pos[i] = ops[i].dest_pos;
        }
        return static_cast<int64_t>(ops.size());
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// The reference's Opcodes of its own editops of one pair (details/types.hpp:526-565):
// five values per opcode (type, src_begin, src_end, dest_begin, dest_end) into out,
// at most 2 * (l1 + l2) + 1 opcodes; returns their count.
int64_t ref_opcodes(int metric, const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, uint64_t* out)
{
    try {
        std::vector<uint32_t> a(s1, s1 + l1), b(s2, s2 + l2);
        rapidfuzz::Editops ops = metric == 0 ? rapidfuzz::levenshtein_editops(a, b) : rapidfuzz::lcs_seq_editops(a, b);
        rapidfuzz::Opcodes oc(ops);
        for (size_t i = 0; i < oc.size(); ++i) {
            out[5 * i] = static_cast<uint64_t>(oc[i].type);
            out[5 * i + 1] = oc[i].src_begin;
            out[5 * i + 2] = oc[i].src_end;
            out[5 * i + 3] = oc[i].dest_begin;
            out[5 * i + 4] = oc[i].dest_end;
        }
        return static_cast<int64_t>(oc.size());
    }
    catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

// CPU baseline: editops of pairs (a[i], b[i]) of 8-bit strings over nthreads threads;
// counts[i] receives pair i's number of operations.
int ref_editops_u8_mt(int metric, const uint8_t* a, const uint64_t* aoff, const uint8_t* b, const uint64_t* boff,
                      size_t n, uint64_t* counts, int nthreads)
{
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        for (;;) {
            size_t i0 = next.fetch_add(256);
            if (i0 >= n) break;
            for (size_t i = i0; i < std::min(n, i0 + 256); ++i) {
                std::basic_string<uint8_t> s1(a + aoff[i], a + aoff[i + 1]), s2(b + boff[i], b + boff[i + 1]);
                rapidfuzz::Editops ops =
                    metric == 0 ? rapidfuzz::levenshtein_editops(s1, s2) : rapidfuzz::lcs_seq_editops(s1, s2);
                counts[i] = ops.size();
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < std::max(1, nthreads); ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
    return 0;
}

// The reference's Editops / Opcodes algebra (details/types.hpp:167-565) on one list of
// operations given as (type, src, dest) triples.  op: 0 slice, 1 remove_slice, 2 reverse,
// 3 inverse, 4 remove_subsequence(sub), 5 Opcodes(ops).slice, 6 Opcodes(ops).inverse,
// 7 Opcodes(ops).reverse, 8 Editops(Opcodes(ops).slice(a, b, c)).  Results: editops as
// triples in out (ops 0-4, 8) or opcodes as 5-tuples (ops 5-7); -1 if the reference throws.
int64_t ref_editops_algebra(int op, const int64_t* ops, size_t n, size_t src_len, size_t dest_len, int a, int b,
                            int c, const int64_t* sub, size_t nsub, int64_t* out, size_t* out_src, size_t* out_dest)
{
    try {
        auto make = [](const int64_t* t, size_t k, size_t sl, size_t dl) {
            rapidfuzz::Editops e;
            for (size_t i = 0; i < k; ++i)
                e.emplace_back(static_cast<rapidfuzz::EditType>(t[3 * i]), static_cast<size_t>(t[3 * i + 1]),
                               static_cast<size_t>(t[3 * i + 2]));
            e.set_src_len(sl);
            e.set_dest_len(dl);
            return e;
        };
        rapidfuzz::Editops e = make(ops, n, src_len, dest_len);
        auto put_ed = [&](const rapidfuzz::Editops& r) {
            for (size_t i = 0; i < r.size(); ++i) {
                out[3 * i] = static_cast<int64_t>(r[i].type);
                out[3 * i + 1] = static_cast<int64_t>(r[i].src_pos);
                out[3 * i + 2] = static_cast<int64_t>(r[i].dest_pos);
            }
            *out_src = r.get_src_len();
            *out_dest = r.get_dest_len();
            return static_cast<int64_t>(r.size());
        };
        auto put_op = [&](const rapidfuzz::Opcodes& r) {
            for (size_t i = 0; i < r.size(); ++i) {
                out[5 * i] = static_cast<int64_t>(r[i].type);
                out[5 * i + 1] = static_cast<int64_t>(r[i].src_begin);
                out[5 * i + 2] = static_cast<int64_t>(r[i].src_end);
                out[5 * i + 3] = static_cast<int64_t>(r[i].dest_begin);
                out[5 * i + 4] = static_cast<int64_t>(r[i].dest_end);
            }
            *out_src = r.get_src_len();
            *out_dest = r.get_dest_len();
            return static_cast<int64_t>(r.size());
        };
        switch (op) {
        case 0: return put_ed(e.slice(a, b, c));
        case 1: e.remove_slice(a, b, c); return put_ed(e);
        case 2: return put_ed(e.reverse());
        case 3: return put_ed(e.inverse());
        case 4: return put_ed(e.remove_subsequence(make(sub, nsub, 0, 0)));
        case 5: return put_op(rapidfuzz::Opcodes(e).slice(a, b, c));
        case 6: return put_op(rapidfuzz::Opcodes(e).inverse());
        case 7: return put_op(rapidfuzz::Opcodes(e).reverse());
        case 8: return put_ed(rapidfuzz::Editops(rapidfuzz::Opcodes(e).slice(a, b, c)));
        default: return -1;
        }
    }
    catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// fuzz::partial_ratio_alignment (fuzz.hpp:129-135) on two strings of 32-bit units:
// *score and al = {src_start, src_end, dest_start, dest_end}.
int ref_partial_ratio_alignment(const uint32_t* s1, size_t l1, const uint32_t* s2, size_t l2, double cutoff,
                                double* score, size_t* al)
{
    try {
        std::basic_string<uint32_t> a(s1, s1 + l1), b(s2, s2 + l2);
        const rapidfuzz::ScoreAlignment<double> r = rapidfuzz::fuzz::partial_ratio_alignment(a, b, cutoff);
        *score = r.score;
        al[0] = r.src_start;
        al[1] = r.src_end;
        al[2] = r.dest_start;
        al[3] = r.dest_end;
        return 0;
    }
    catch (const std::exception& ex) {
        g_err = ex.what();
        return -1;
    }
}

// partial_ratio_alignment over 8-bit pairs on nthreads threads: score[i], al[4 i .. 4 i + 3]
int ref_partial_ratio_alignment_u8_mt(const uint8_t* a, const uint64_t* aoff, const uint8_t* b, const uint64_t* boff,
                                      size_t n, double cutoff, double* score, uint64_t* al, int nthreads)
{
    std::atomic<size_t> next(0);
    auto worker = [&]() {
        for (;;) {
            size_t i0 = next.fetch_add(64);
            if (i0 >= n) break;
            for (size_t i = i0; i < std::min(n, i0 + 64); ++i) {
                std::basic_string<uint8_t> s1(a + aoff[i], a + aoff[i + 1]), s2(b + boff[i], b + boff[i + 1]);
                const rapidfuzz::ScoreAlignment<double> r = rapidfuzz::fuzz::partial_ratio_alignment(s1, s2, cutoff);
                score[i] = r.score;
                al[4 * i] = r.src_start;
                al[4 * i + 1] = r.src_end;
                al[4 * i + 2] = r.dest_start;
                al[4 * i + 3] = r.dest_end;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < std::max(1, nthreads); ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
    return 0;
}

} // extern "C"
