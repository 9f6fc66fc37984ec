// rapidfuzz_amd.hpp -- C++17 drop-in for the reference's scorer classes,
// running on one MI355X through the C ABI (include/rapidfuzz_amd.h).
//
// The classes keep the reference's names, constructors, method signatures and
// defaults:
//   CachedLevenshtein (distance/Levenshtein.hpp:421-492), CachedIndel
//   (distance/Indel.hpp:148-188), CachedLCSseq (distance/LCSseq.hpp:197-223),
//   CachedOSA (distance/OSA.hpp:231-275), CachedJaro (distance/Jaro.hpp:195-230),
//   CachedJaroWinkler (distance/JaroWinkler.hpp:168-209), CachedHamming
//   (distance/Hamming.hpp:131-161), CachedPrefix (distance/Prefix.hpp:65-94),
//   CachedPostfix (distance/Postfix.hpp:65-94), fuzz::CachedRatio
//   (fuzz.hpp:98-117), experimental::Multi{Levenshtein,Indel,LCSseq,OSA,Jaro,
//   JaroWinkler}<MaxLen> and fuzz::experimental::MultiRatio<MaxLen>
//   (e.g. Levenshtein.hpp:297-419, details/distance.hpp:390-582).
// A one-pair call is one device round trip; the batch methods added here
// (distance_many and friends over a resident `Choices` set, extract_topk)
// are what the GPU is for.
//
// Errors: std::invalid_argument where the reference throws it (same messages
// for the Multi* checks), rapidfuzz_amd::Error for device failures.
// Threading: const methods may be called from several threads (calls on one
// device context are serialized inside the library).
//
// Code units: each character is converted to an unsigned code unit
// (char -> uint8_t, char16_t -> uint16_t, char32_t / wchar_t -> uint32_t).
// When query and choice widths differ both are widened to the larger.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/rapidfuzz_amd.h"

namespace rapidfuzz_amd {

// LevenshteinWeightTable (details/types.hpp:19-23)
struct LevenshteinWeightTable {
    size_t insert_cost;
    size_t delete_cost;
    size_t replace_cost;
};

struct Error : std::runtime_error {
    rfa_status status;
    Error(rfa_status st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

namespace detail {

inline void check(rfa_status st)
{
    if (st == RFA_OK) return;
    const std::string msg = rfa_last_error();
    if (st == RFA_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw Error(st, msg);
}

// Process-wide context per device, created on first use.
inline rfa_context* context(int device = 0)
{
    static std::mutex mu;
    static std::map<int, std::shared_ptr<rfa_context>> ctxs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second.get();
    rfa_context* c = nullptr;
    check(rfa_context_create(device, &c));
    ctxs[device] = std::shared_ptr<rfa_context>(c, [](rfa_context* p) { rfa_context_destroy(p); });
    return c;
}

template <typename T>
uint32_t to_unit(T c)
{
    if constexpr (std::is_integral<T>::value)
        return static_cast<uint32_t>(static_cast<typename std::make_unsigned<T>::type>(c));
    else
        return static_cast<uint32_t>(c);
}

template <typename T>
constexpr int width_of()
{
    return sizeof(T) == 1 ? 1 : (sizeof(T) == 2 ? 2 : 4);
}

// Sentences: anything with begin()/end(); character pointers and character
// arrays (string literals) are C strings that end at the first NUL.
template <typename S>
auto to_begin(const S& s)
{
    if constexpr (std::is_array<S>::value)
        return &s[0];
    else if constexpr (std::is_pointer<S>::value)
        return s;
    else
        return std::begin(s);
}
template <typename S>
auto to_end(const S& s)
{
    if constexpr (std::is_array<S>::value) {
        size_t n = 0;
        while (n < std::extent<S>::value && s[n]) ++n;
        return &s[0] + n;
    }
    else if constexpr (std::is_pointer<S>::value) {
        auto e = s;
        while (*e) ++e;
        return e;
    }
    else
        return std::end(s);
}

template <typename S>
using sentence_char_t = typename std::decay<decltype(*to_begin(std::declval<const S&>()))>::type;

// Packed code units of a list of strings (kept as 32-bit units plus the widest width seen).
struct Packed {
    std::vector<uint32_t> units;
    std::vector<uint64_t> off{0};
    int width = 1;

    template <typename It>
    void add(It first, It last)
    {
        using T = typename std::decay<decltype(*first)>::type;
        width = std::max(width, width_of<T>());
        for (; first != last; ++first) units.push_back(to_unit(*first));
        off.push_back(units.size());
    }
    size_t count() const { return off.size() - 1; }
};

// A device string set of a given width built from packed units.
class DeviceStrings {
public:
    DeviceStrings(rfa_context* ctx, const Packed& p, int width)
    {
        const size_t n = p.units.size();
        std::vector<uint8_t> bytes(std::max<size_t>(n * width, 1));
        for (size_t i = 0; i < n; ++i) {
            const uint32_t u = p.units[i];
            if (width == 1)
                bytes[i] = static_cast<uint8_t>(u);
            else if (width == 2) {
                const uint16_t v = static_cast<uint16_t>(u);
                std::memcpy(&bytes[i * 2], &v, 2);
            }
            else
                std::memcpy(&bytes[i * 4], &u, 4);
        }
        check(rfa_strings_create(ctx, bytes.data(), width, p.off.data(), p.count(), &h_));
    }
    ~DeviceStrings() { rfa_strings_destroy(h_); }
    DeviceStrings(const DeviceStrings&) = delete;
    DeviceStrings& operator=(const DeviceStrings&) = delete;
    rfa_strings* get() const { return h_; }

private:
    rfa_strings* h_ = nullptr;
};

// rfa_scorer over packed queries, one per code-unit width (built on demand).
class ScorerSet {
public:
    // signed8: 8-bit units order as signed values when tokens are sorted (`char` queries)
    ScorerSet(int metric, Packed q, rfa_weights w, double pw, bool signed8 = false)
        : metric_(metric), q_(std::move(q)), w_(w), pw_(pw), signed8_(signed8)
    {}
    ~ScorerSet()
    {
        for (auto& kv : scorers_) rfa_scorer_destroy(kv.second);
    }
    ScorerSet(const ScorerSet&) = delete;
    ScorerSet& operator=(const ScorerSet&) = delete;

    int query_width() const { return q_.width; }
    size_t count() const { return q_.count(); }

    rfa_scorer* get(int width) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = scorers_.find(width);
        if (it != scorers_.end()) return it->second;
        rfa_context* ctx = context();
        DeviceStrings qs(ctx, q_, width);
        rfa_scorer* sc = nullptr;
        check(rfa_scorer_create(ctx, metric_, qs.get(), &w_, pw_, &sc));
        if (signed8_ && width == 1) {
            const rfa_status st = rfa_scorer_set_option(sc, RFA_SCORER_SIGNED_UNITS, 1);
            if (st != RFA_OK) {
                rfa_scorer_destroy(sc);
                check(st);
            }
        }
        scorers_[width] = sc;
        return sc;
    }

private:
    int metric_;
    Packed q_;
    rfa_weights w_;
    double pw_;
    bool signed8_;
    mutable std::mutex mu_;
    mutable std::map<int, rfa_scorer*> scorers_;
};

template <typename R>
constexpr bool is_int_result()
{
    return std::is_integral<R>::value;
}

} // namespace detail

// A device-resident set of choices, uploaded once and scored many times.
class Choices {
public:
    template <typename It>
    Choices(It first, It last)
    {
        for (; first != last; ++first) p_.add(detail::to_begin(*first), detail::to_end(*first));
    }
    template <typename Container>
    explicit Choices(const Container& c) : Choices(std::begin(c), std::end(c))
    {}
    size_t size() const { return p_.count(); }
    int width() const { return p_.width; }
    size_t length(size_t i) const { return p_.off[i + 1] - p_.off[i]; }

    // the set at a code-unit width (uploaded on first use)
    rfa_strings* at_width(int w) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = sets_.find(w);
        if (it != sets_.end()) return it->second->get();
        auto s = std::make_shared<detail::DeviceStrings>(detail::context(), p_, w);
        sets_[w] = s;
        return s->get();
    }

private:
    detail::Packed p_;
    mutable std::mutex mu_;
    mutable std::map<int, std::shared_ptr<detail::DeviceStrings>> sets_;
};

namespace detail {

// Shared implementation of the Cached* scorers: one query, ResType results.
template <int Metric, typename ResType, int64_t WorstSimilarity, int64_t WorstDistance>
class CachedBase {
public:
    size_t query_size() const { return len_; }

protected:
    template <typename It>
    void init(It first, It last, rfa_weights w, double pw)
    {
        Packed q;
        q.add(first, last);
        len_ = q.count() ? q.off[1] : 0;
        set_ = std::make_shared<ScorerSet>(Metric, std::move(q), w, pw);
    }

    // Hamming without padding: unequal lengths are the reference's error (Hamming_impl.hpp:27)
    void check_length(size_t len2) const
    {
        if (equal_len_ && len2 != len_) throw std::invalid_argument("Sequences are not the same length.");
    }

    // scores of the query against `choices`; out has choices.size() entries
    template <typename Out>
    void run(const Choices& choices, int kind, Out* out, uint64_t cu, double cd, int semantics = RFA_CACHED) const
    {
        if (!choices.size()) return;
        if (equal_len_)
            for (size_t i = 0; i < choices.size(); ++i) check_length(choices.length(i));
        const int w = std::max(set_->query_width(), choices.width());
        check(rfa_cdist(set_->get(w), choices.at_width(w), kind, semantics, cu, cd, out, 0));
    }

    template <typename It, typename Out>
    Out one(It first, It last, int kind, uint64_t cu, double cd) const
    {
        Packed p;
        p.add(first, last);
        check_length(p.off[1]);
        const int w = std::max(set_->query_width(), p.width);
        DeviceStrings s(context(), p, w);
        Out v{};
        check(rfa_cdist(set_->get(w), s.get(), kind, RFA_CACHED, cu, cd, &v, 0));
        return v;
    }

    static uint64_t u_cut(ResType c) { return static_cast<uint64_t>(c); }

    std::shared_ptr<ScorerSet> set_;
    size_t len_ = 0;
    bool equal_len_ = false;

public:
    // ---- the reference's methods (details/distance.hpp:211-301, 326-356) ----
    template <typename InputIt2>
    ResType distance(InputIt2 first2, InputIt2 last2, ResType score_cutoff = static_cast<ResType>(WorstDistance),
                     ResType score_hint = static_cast<ResType>(WorstDistance)) const
    {
        (void)score_hint;
        if constexpr (is_int_result<ResType>())
            return static_cast<ResType>(one<InputIt2, uint64_t>(first2, last2, RFA_DISTANCE, u_cut(score_cutoff), 0.0));
        else
            return one<InputIt2, double>(first2, last2, RFA_DISTANCE, 0, score_cutoff);
    }
    template <typename Sentence2>
    ResType distance(const Sentence2& s2, ResType score_cutoff = static_cast<ResType>(WorstDistance),
                     ResType score_hint = static_cast<ResType>(WorstDistance)) const
    {
        return distance(to_begin(s2), to_end(s2), score_cutoff, score_hint);
    }

    template <typename InputIt2>
    ResType similarity(InputIt2 first2, InputIt2 last2, ResType score_cutoff = static_cast<ResType>(WorstSimilarity),
                       ResType score_hint = static_cast<ResType>(WorstSimilarity)) const
    {
        (void)score_hint;
        if constexpr (is_int_result<ResType>())
            return static_cast<ResType>(
                one<InputIt2, uint64_t>(first2, last2, RFA_SIMILARITY, u_cut(score_cutoff), 0.0));
        else
            return one<InputIt2, double>(first2, last2, RFA_SIMILARITY, 0, score_cutoff);
    }
    template <typename Sentence2>
    ResType similarity(const Sentence2& s2, ResType score_cutoff = static_cast<ResType>(WorstSimilarity),
                       ResType score_hint = static_cast<ResType>(WorstSimilarity)) const
    {
        return similarity(to_begin(s2), to_end(s2), score_cutoff, score_hint);
    }

    template <typename InputIt2>
    double normalized_distance(InputIt2 first2, InputIt2 last2, double score_cutoff = 1.0,
                               double score_hint = 1.0) const
    {
        (void)score_hint;
        return one<InputIt2, double>(first2, last2, RFA_NORMALIZED_DISTANCE, 0, score_cutoff);
    }
    template <typename Sentence2>
    double normalized_distance(const Sentence2& s2, double score_cutoff = 1.0, double score_hint = 1.0) const
    {
        return normalized_distance(to_begin(s2), to_end(s2), score_cutoff, score_hint);
    }

    template <typename InputIt2>
    double normalized_similarity(InputIt2 first2, InputIt2 last2, double score_cutoff = 0.0,
                                 double score_hint = 0.0) const
    {
        (void)score_hint;
        return one<InputIt2, double>(first2, last2, RFA_NORMALIZED_SIMILARITY, 0, score_cutoff);
    }
    template <typename Sentence2>
    double normalized_similarity(const Sentence2& s2, double score_cutoff = 0.0, double score_hint = 0.0) const
    {
        return normalized_similarity(to_begin(s2), to_end(s2), score_cutoff, score_hint);
    }

    // ---- batch: the query against every choice (one launch) ----
    std::vector<ResType> distance_many(const Choices& choices,
                                       ResType score_cutoff = static_cast<ResType>(WorstDistance)) const
    {
        return many(choices, RFA_DISTANCE, score_cutoff);
    }
    std::vector<ResType> similarity_many(const Choices& choices,
                                         ResType score_cutoff = static_cast<ResType>(WorstSimilarity)) const
    {
        return many(choices, RFA_SIMILARITY, score_cutoff);
    }
    std::vector<double> normalized_distance_many(const Choices& choices, double score_cutoff = 1.0) const
    {
        std::vector<double> out(choices.size());
        run(choices, RFA_NORMALIZED_DISTANCE, out.data(), 0, score_cutoff);
        return out;
    }
    std::vector<double> normalized_similarity_many(const Choices& choices, double score_cutoff = 0.0) const
    {
        std::vector<double> out(choices.size());
        run(choices, RFA_NORMALIZED_SIMILARITY, out.data(), 0, score_cutoff);
        return out;
    }

private:
    std::vector<ResType> many(const Choices& choices, int kind, ResType cutoff) const
    {
        std::vector<ResType> out(choices.size());
        if constexpr (is_int_result<ResType>()) {
            std::vector<uint64_t> tmp(choices.size());
            run(choices, kind, tmp.data(), u_cut(cutoff), 0.0);
            std::transform(tmp.begin(), tmp.end(), out.begin(), [](uint64_t v) { return static_cast<ResType>(v); });
        }
        else {
            run(choices, kind, out.data(), 0, cutoff);
        }
        return out;
    }
};

constexpr int64_t kIntMax = std::numeric_limits<int64_t>::max();

} // namespace detail

template <typename CharT1>
struct CachedLevenshtein : detail::CachedBase<RFA_LEVENSHTEIN, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedLevenshtein(const Sentence1& s1_, LevenshteinWeightTable aWeights = {1, 1, 1})
        : CachedLevenshtein(detail::to_begin(s1_), detail::to_end(s1_), aWeights)
    {}
    template <typename InputIt1>
    CachedLevenshtein(InputIt1 first1, InputIt1 last1, LevenshteinWeightTable aWeights = {1, 1, 1})
    {
        this->init(first1, last1, {aWeights.insert_cost, aWeights.delete_cost, aWeights.replace_cost}, 0.1);
    }
};

template <typename CharT1>
struct CachedIndel : detail::CachedBase<RFA_INDEL, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedIndel(const Sentence1& s1_) : CachedIndel(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedIndel(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

template <typename CharT1>
struct CachedLCSseq : detail::CachedBase<RFA_LCSSEQ, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedLCSseq(const Sentence1& s1_) : CachedLCSseq(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedLCSseq(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

template <typename CharT1>
struct CachedOSA : detail::CachedBase<RFA_OSA, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedOSA(const Sentence1& s1_) : CachedOSA(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedOSA(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

template <typename CharT1>
struct CachedJaro : detail::CachedBase<RFA_JARO, double, 0, 1> {
    template <typename Sentence1>
    explicit CachedJaro(const Sentence1& s1_) : CachedJaro(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedJaro(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

template <typename CharT1>
struct CachedJaroWinkler : detail::CachedBase<RFA_JARO_WINKLER, double, 0, 1> {
    template <typename Sentence1>
    explicit CachedJaroWinkler(const Sentence1& s1_, double _prefix_weight = 0.1)
        : CachedJaroWinkler(detail::to_begin(s1_), detail::to_end(s1_), _prefix_weight)
    {}
    template <typename InputIt1>
    CachedJaroWinkler(InputIt1 first1, InputIt1 last1, double _prefix_weight = 0.1)
    {
        this->init(first1, last1, {1, 1, 1}, _prefix_weight);
    }
};

// Hamming.hpp:131-161 (pad = false: unequal lengths throw std::invalid_argument)
template <typename CharT1>
struct CachedHamming : detail::CachedBase<RFA_HAMMING, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedHamming(const Sentence1& s1_, bool pad_ = true)
        : CachedHamming(detail::to_begin(s1_), detail::to_end(s1_), pad_)
    {}
    template <typename InputIt1>
    CachedHamming(InputIt1 first1, InputIt1 last1, bool pad_ = true)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
        this->equal_len_ = !pad_;
    }
};

// Prefix.hpp:65-94
template <typename CharT1>
struct CachedPrefix : detail::CachedBase<RFA_PREFIX, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedPrefix(const Sentence1& s1_) : CachedPrefix(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedPrefix(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

// Postfix.hpp:65-94
template <typename CharT1>
struct CachedPostfix : detail::CachedBase<RFA_POSTFIX, size_t, 0, detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedPostfix(const Sentence1& s1_) : CachedPostfix(detail::to_begin(s1_), detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedPostfix(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

// Deduction guides (e.g. Levenshtein.hpp:484-492)
template <typename Sentence1>
explicit CachedLevenshtein(const Sentence1& s1_, LevenshteinWeightTable aWeights = {1, 1, 1})
    -> CachedLevenshtein<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedIndel(const Sentence1& s1_) -> CachedIndel<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedLCSseq(const Sentence1& s1_) -> CachedLCSseq<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedOSA(const Sentence1& s1_) -> CachedOSA<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedJaro(const Sentence1& s1_) -> CachedJaro<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedJaroWinkler(const Sentence1& s1_, double _prefix_weight = 0.1)
    -> CachedJaroWinkler<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedHamming(const Sentence1& s1_, bool pad_ = true) -> CachedHamming<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedPrefix(const Sentence1& s1_) -> CachedPrefix<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
explicit CachedPostfix(const Sentence1& s1_) -> CachedPostfix<detail::sentence_char_t<Sentence1>>;

// ---- free functions -----------------------------------------------------------
// levenshtein_* (Levenshtein.hpp:140-265), indel_* (Indel.hpp:12-60), lcs_seq_*
// (LCSseq.hpp:12-60), osa_* (OSA.hpp:36-108), jaro_* (Jaro.hpp:12-59),
// jaro_winkler_* (JaroWinkler.hpp:12-75), hamming_* (Hamming.hpp:35-125),
// prefix_* / postfix_* (Prefix.hpp:12-60, Postfix.hpp:12-60): the same names,
// parameters and defaults.  Each call scores its one pair through a one-query
// scorer, so the result is the Cached* scorer's (the reference's free functions
// and Cached* classes share details/distance.hpp's normalization); a call is a
// device round trip, and many pairs belong in one pairs_*() call.
namespace detail {
template <typename It>
using iter_char_t = typename std::decay<decltype(*std::declval<It>())>::type;
} // namespace detail

#define RFA_FREE_KINDS(NAME, CACHED, DECL_EXTRA, CTOR_EXTRA)                                                        \
    template <typename InputIt1, typename InputIt2>                                                                \
    size_t NAME##_distance(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2 DECL_EXTRA,           \
                           size_t score_cutoff = std::numeric_limits<size_t>::max())                               \
    {                                                                                                              \
        return CACHED<detail::iter_char_t<InputIt1>>(first1, last1 CTOR_EXTRA).distance(first2, last2, score_cutoff); \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    size_t NAME##_distance(const Sentence1& s1, const Sentence2& s2 DECL_EXTRA,                                   \
                           size_t score_cutoff = std::numeric_limits<size_t>::max())                               \
    {                                                                                                              \
        return NAME##_distance(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2)  \
                                   CTOR_EXTRA, score_cutoff);                                                      \
    }                                                                                                              \
    template <typename InputIt1, typename InputIt2>                                                                \
    size_t NAME##_similarity(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2 DECL_EXTRA,         \
                             size_t score_cutoff = 0)                                                              \
    {                                                                                                              \
        return CACHED<detail::iter_char_t<InputIt1>>(first1, last1 CTOR_EXTRA)                                    \
            .similarity(first2, last2, score_cutoff);                                                              \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    size_t NAME##_similarity(const Sentence1& s1, const Sentence2& s2 DECL_EXTRA, size_t score_cutoff = 0)         \
    {                                                                                                              \
        return NAME##_similarity(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),                   \
                                 detail::to_end(s2) CTOR_EXTRA, score_cutoff);                                     \
    }                                                                                                              \
    template <typename InputIt1, typename InputIt2>                                                                \
    double NAME##_normalized_distance(InputIt1 first1, InputIt1 last1, InputIt2 first2,                            \
                                      InputIt2 last2 DECL_EXTRA, double score_cutoff = 1.0)                        \
    {                                                                                                              \
        return CACHED<detail::iter_char_t<InputIt1>>(first1, last1 CTOR_EXTRA)                                    \
            .normalized_distance(first2, last2, score_cutoff);                                                     \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    double NAME##_normalized_distance(const Sentence1& s1, const Sentence2& s2 DECL_EXTRA,                         \
                                      double score_cutoff = 1.0)                                                   \
    {                                                                                                              \
        return NAME##_normalized_distance(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),          \
                                          detail::to_end(s2) CTOR_EXTRA, score_cutoff);                            \
    }                                                                                                              \
    template <typename InputIt1, typename InputIt2>                                                                \
    double NAME##_normalized_similarity(InputIt1 first1, InputIt1 last1, InputIt2 first2,                          \
                                        InputIt2 last2 DECL_EXTRA, double score_cutoff = 0.0)                      \
    {                                                                                                              \
        return CACHED<detail::iter_char_t<InputIt1>>(first1, last1 CTOR_EXTRA)                                    \
            .normalized_similarity(first2, last2, score_cutoff);                                                   \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    double NAME##_normalized_similarity(const Sentence1& s1, const Sentence2& s2 DECL_EXTRA,                       \
                                        double score_cutoff = 0.0)                                                 \
    {                                                                                                              \
        return NAME##_normalized_similarity(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),        \
                                            detail::to_end(s2) CTOR_EXTRA, score_cutoff);                          \
    }

#define RFA_NOTHING
#define RFA_COMMA_PAD_DECL , bool pad_ = true
#define RFA_COMMA_PAD , pad_
RFA_FREE_KINDS(indel, CachedIndel, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_KINDS(lcs_seq, CachedLCSseq, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_KINDS(osa, CachedOSA, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_KINDS(prefix, CachedPrefix, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_KINDS(postfix, CachedPostfix, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_KINDS(hamming, CachedHamming, RFA_COMMA_PAD_DECL, RFA_COMMA_PAD)
#undef RFA_FREE_KINDS
#undef RFA_COMMA_PAD_DECL
#undef RFA_COMMA_PAD

// levenshtein_* carry the weight table and score_hint (Levenshtein.hpp:140-265)
template <typename InputIt1, typename InputIt2>
size_t levenshtein_distance(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                            LevenshteinWeightTable weights = {1, 1, 1},
                            size_t score_cutoff = std::numeric_limits<size_t>::max(),
                            size_t score_hint = std::numeric_limits<size_t>::max())
{
    return CachedLevenshtein<detail::iter_char_t<InputIt1>>(first1, last1, weights)
        .distance(first2, last2, score_cutoff, score_hint);
}
template <typename Sentence1, typename Sentence2>
size_t levenshtein_distance(const Sentence1& s1, const Sentence2& s2, LevenshteinWeightTable weights = {1, 1, 1},
                            size_t score_cutoff = std::numeric_limits<size_t>::max(),
                            size_t score_hint = std::numeric_limits<size_t>::max())
{
    return levenshtein_distance(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2),
                                weights, score_cutoff, score_hint);
}
template <typename InputIt1, typename InputIt2>
size_t levenshtein_similarity(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                              LevenshteinWeightTable weights = {1, 1, 1}, size_t score_cutoff = 0,
                              size_t score_hint = 0)
{
    return CachedLevenshtein<detail::iter_char_t<InputIt1>>(first1, last1, weights)
        .similarity(first2, last2, score_cutoff, score_hint);
}
template <typename Sentence1, typename Sentence2>
size_t levenshtein_similarity(const Sentence1& s1, const Sentence2& s2, LevenshteinWeightTable weights = {1, 1, 1},
                              size_t score_cutoff = 0, size_t score_hint = 0)
{
    return levenshtein_similarity(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2),
                                  weights, score_cutoff, score_hint);
}
template <typename InputIt1, typename InputIt2>
double levenshtein_normalized_distance(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                                       LevenshteinWeightTable weights = {1, 1, 1}, double score_cutoff = 1.0,
                                       double score_hint = 1.0)
{
    return CachedLevenshtein<detail::iter_char_t<InputIt1>>(first1, last1, weights)
        .normalized_distance(first2, last2, score_cutoff, score_hint);
}
template <typename Sentence1, typename Sentence2>
double levenshtein_normalized_distance(const Sentence1& s1, const Sentence2& s2,
                                       LevenshteinWeightTable weights = {1, 1, 1}, double score_cutoff = 1.0,
                                       double score_hint = 1.0)
{
    return levenshtein_normalized_distance(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),
                                           detail::to_end(s2), weights, score_cutoff, score_hint);
}
template <typename InputIt1, typename InputIt2>
double levenshtein_normalized_similarity(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                                         LevenshteinWeightTable weights = {1, 1, 1}, double score_cutoff = 0.0,
                                         double score_hint = 0.0)
{
    return CachedLevenshtein<detail::iter_char_t<InputIt1>>(first1, last1, weights)
        .normalized_similarity(first2, last2, score_cutoff, score_hint);
}
template <typename Sentence1, typename Sentence2>
double levenshtein_normalized_similarity(const Sentence1& s1, const Sentence2& s2,
                                         LevenshteinWeightTable weights = {1, 1, 1}, double score_cutoff = 0.0,
                                         double score_hint = 0.0)
{
    return levenshtein_normalized_similarity(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),
                                             detail::to_end(s2), weights, score_cutoff, score_hint);
}

// jaro_* and jaro_winkler_* return doubles for every kind (Jaro.hpp:12-59, JaroWinkler.hpp:12-75)
#define RFA_FREE_DOUBLE(NAME, KIND, CUT, CACHED, DECL_EXTRA, CTOR_EXTRA)                                            \
    template <typename InputIt1, typename InputIt2>                                                                \
    double NAME##_##KIND(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2 DECL_EXTRA,             \
                         double score_cutoff = CUT)                                                                \
    {                                                                                                              \
        return CACHED<detail::iter_char_t<InputIt1>>(first1, last1 CTOR_EXTRA).KIND(first2, last2, score_cutoff);   \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    double NAME##_##KIND(const Sentence1& s1, const Sentence2& s2 DECL_EXTRA, double score_cutoff = CUT)          \
    {                                                                                                              \
        return NAME##_##KIND(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2),                       \
                             detail::to_end(s2) CTOR_EXTRA, score_cutoff);                                         \
    }
#define RFA_COMMA_PW_DECL , double prefix_weight = 0.1
#define RFA_COMMA_PW , prefix_weight
RFA_FREE_DOUBLE(jaro, distance, 1.0, CachedJaro, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_DOUBLE(jaro, similarity, 0.0, CachedJaro, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_DOUBLE(jaro, normalized_distance, 1.0, CachedJaro, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_DOUBLE(jaro, normalized_similarity, 0.0, CachedJaro, RFA_NOTHING, RFA_NOTHING)
RFA_FREE_DOUBLE(jaro_winkler, distance, 1.0, CachedJaroWinkler, RFA_COMMA_PW_DECL, RFA_COMMA_PW)
RFA_FREE_DOUBLE(jaro_winkler, similarity, 0.0, CachedJaroWinkler, RFA_COMMA_PW_DECL, RFA_COMMA_PW)
RFA_FREE_DOUBLE(jaro_winkler, normalized_distance, 1.0, CachedJaroWinkler, RFA_COMMA_PW_DECL, RFA_COMMA_PW)
RFA_FREE_DOUBLE(jaro_winkler, normalized_similarity, 0.0, CachedJaroWinkler, RFA_COMMA_PW_DECL, RFA_COMMA_PW)
#undef RFA_COMMA_PW_DECL
#undef RFA_COMMA_PW

// ---- the windows of a partial match (details/types.hpp:567-592) ----
template <typename T>
struct ScoreAlignment {
    T score = T();
    size_t src_start = 0; // window of the first string
    size_t src_end = 0;
    size_t dest_start = 0; // window of the second string
    size_t dest_end = 0;
    ScoreAlignment() = default;
    ScoreAlignment(T s, size_t ss, size_t se, size_t ds, size_t de)
        : score(s), src_start(ss), src_end(se), dest_start(ds), dest_end(de)
    {}
    friend bool operator==(const ScoreAlignment& a, const ScoreAlignment& b)
    {
        return a.score == b.score && a.src_start == b.src_start && a.src_end == b.src_end &&
               a.dest_start == b.dest_start && a.dest_end == b.dest_end;
    }
};

// ---- edit operations (details/types.hpp:28-60, 167-355: the parts a caller reads) ----
enum class EditType { None = RFA_EDIT_NONE, Replace = RFA_EDIT_REPLACE, Insert = RFA_EDIT_INSERT, Delete = RFA_EDIT_DELETE };

struct EditOp {
    EditType type = EditType::None;
    size_t src_pos = 0;
    size_t dest_pos = 0;
    EditOp() = default;
    EditOp(EditType t, size_t s, size_t d) : type(t), src_pos(s), dest_pos(d) {}
    friend bool operator==(const EditOp& a, const EditOp& b)
    {
        return a.type == b.type && a.src_pos == b.src_pos && a.dest_pos == b.dest_pos;
    }
    friend bool operator!=(const EditOp& a, const EditOp& b) { return !(a == b); }
};

namespace detail {
// Python-style slice bounds of details/types.hpp:108-160: negative indices count from
// the end, indices are clamped to [0, size]; steps must be positive.  Returns false for
// an empty range.
inline bool slice_range(size_t size, int& start, int& stop, int step)
{
    if (step == 0) throw std::invalid_argument("slice step cannot be zero");
    if (step < 0) throw std::invalid_argument("step sizes below 0 lead to an invalid order of editops");
    const int n = static_cast<int>(size);
    auto clamp = [n](int i) { return i < 0 ? std::max(i + n, 0) : std::min(i, n); };
    start = clamp(start);
    stop = clamp(stop);
    return start < stop;
}

// the edit-operation lists: a vector with the two string lengths and the reference's slicing
template <typename Op, typename Self>
class OpList : private std::vector<Op> {
    using Base = std::vector<Op>;

public:
    using typename Base::size_type;
    OpList() = default;
    OpList(size_type count, const Op& value) : Base(count, value) {}
    explicit OpList(size_type count) : Base(count) {}

    using Base::at;
    using Base::back;
    using Base::data;
    using Base::front;
    using Base::operator[];
    using Base::begin;
    using Base::cbegin;
    using Base::cend;
    using Base::crbegin;
    using Base::crend;
    using Base::end;
    using Base::rbegin;
    using Base::rend;
    using Base::capacity;
    using Base::empty;
    using Base::max_size;
    using Base::reserve;
    using Base::shrink_to_fit;
    using Base::size;
    using Base::clear;
    using Base::emplace;
    using Base::emplace_back;
    using Base::erase;
    using Base::insert;
    using Base::pop_back;
    using Base::push_back;
    using Base::resize;

    size_t get_src_len() const noexcept { return src_len_; }
    size_t get_dest_len() const noexcept { return dest_len_; }
    void set_src_len(size_t n) noexcept { src_len_ = n; }
    void set_dest_len(size_t n) noexcept { dest_len_ = n; }
    void swap(Self& other) noexcept
    {
        std::swap(src_len_, other.src_len_);
        std::swap(dest_len_, other.dest_len_);
        Base::swap(static_cast<Base&>(other));
    }
    // elements start, start + step, ... below stop; the lengths are kept
    Self slice(int start, int stop, int step = 1) const
    {
        Self out;
        out.set_src_len(src_len_);
        out.set_dest_len(dest_len_);
        if (!slice_range(size(), start, stop, step)) return out;
        out.reserve(static_cast<size_t>((stop - 1 - start) / step + 1));
        for (int i = start; i < stop; i += step) out.push_back((*this)[static_cast<size_t>(i)]);
        return out;
    }
    // the same elements removed
    void remove_slice(int start, int stop, int step = 1)
    {
        if (!slice_range(size(), start, stop, step)) return;
        size_t w = static_cast<size_t>(start);
        for (size_t i = static_cast<size_t>(start); i < size(); ++i) {
            const int k = static_cast<int>(i);
            if (k >= stop || (k - start) % step != 0) (*this)[w++] = (*this)[i];
        }
        resize(w);
        shrink_to_fit();
    }
    Self reverse() const
    {
        Self out = static_cast<const Self&>(*this);
        std::reverse(out.begin(), out.end());
        return out;
    }
    friend bool operator==(const Self& a, const Self& b)
    {
        return a.src_len_ == b.src_len_ && a.dest_len_ == b.dest_len_ &&
               static_cast<const Base&>(a) == static_cast<const Base&>(b);
    }
    friend bool operator!=(const Self& a, const Self& b) { return !(a == b); }
    friend void swap(Self& a, Self& b) noexcept { a.swap(b); }

private:
    size_t src_len_ = 0, dest_len_ = 0;
};
} // namespace detail

class Opcodes;

// details/types.hpp:167-355
class Editops : public detail::OpList<EditOp, Editops> {
    using List = detail::OpList<EditOp, Editops>;

public:
    using List::List;
    Editops() = default;
    Editops(const Opcodes& ops);
    // insertions and deletions exchanged, source and destination swapped
    Editops inverse() const
    {
        Editops out = *this;
        out.set_src_len(get_dest_len());
        out.set_dest_len(get_src_len());
        for (EditOp& op : out) {
            std::swap(op.src_pos, op.dest_pos);
            if (op.type == EditType::Delete)
                op.type = EditType::Insert;
            else if (op.type == EditType::Insert)
                op.type = EditType::Delete;
        }
        return out;
    }
    // the operations of `sub` (a subsequence, in order) taken out; the source positions of
    // the remaining ones move by the insertions (+1) and deletions (-1) removed before them
    Editops remove_subsequence(const Editops& sub) const
    {
        if (sub.size() > size()) throw std::invalid_argument("subsequence is not a subsequence");
        Editops out;
        out.set_src_len(get_src_len());
        out.set_dest_len(get_dest_len());
        out.reserve(size() - sub.size());
        ptrdiff_t shift = 0;
        size_t i = 0;
        auto keep = [&](const EditOp& op) {
            EditOp o = op;
            o.src_pos = static_cast<size_t>(static_cast<ptrdiff_t>(o.src_pos) + shift);
            out.push_back(o);
        };
        for (const EditOp& s : sub) {
            while (i < size() && (*this)[i] != s) keep((*this)[i++]);
            if (i == size()) throw std::invalid_argument("subsequence is not a subsequence");
            if (s.type == EditType::Insert)
                ++shift;
            else if (s.type == EditType::Delete)
                --shift;
            ++i;
        }
        while (i < size()) keep((*this)[i++]);
        return out;
    }
};

// details/types.hpp:81-100, 357-565: runs of edit operations with their source and
// destination ranges (EditType::None for the unchanged runs between them)
struct Opcode {
    EditType type = EditType::None;
    size_t src_begin = 0, src_end = 0, dest_begin = 0, dest_end = 0;
    Opcode() = default;
    Opcode(EditType t, size_t sb, size_t se, size_t db, size_t de)
        : type(t), src_begin(sb), src_end(se), dest_begin(db), dest_end(de)
    {}
    friend bool operator==(const Opcode& a, const Opcode& b)
    {
        return a.type == b.type && a.src_begin == b.src_begin && a.src_end == b.src_end &&
               a.dest_begin == b.dest_begin && a.dest_end == b.dest_end;
    }
    friend bool operator!=(const Opcode& a, const Opcode& b) { return !(a == b); }
};

class Opcodes : public detail::OpList<Opcode, Opcodes> {
    using List = detail::OpList<Opcode, Opcodes>;

public:
    using List::List;
    Opcodes() = default;
    // Editops -> Opcodes: consecutive operations of one type that continue each other
    // become one run; the stretches between runs (and after the last) are None runs
    Opcodes(const Editops& ops)
    {
        set_src_len(ops.get_src_len());
        set_dest_len(ops.get_dest_len());
        size_t sp = 0, dp = 0, i = 0;
        while (i < ops.size()) {
            if (sp < ops[i].src_pos || dp < ops[i].dest_pos) {
                push_back(Opcode(EditType::None, sp, ops[i].src_pos, dp, ops[i].dest_pos));
                sp = ops[i].src_pos;
                dp = ops[i].dest_pos;
            }
            const size_t sb = sp, db = dp;
            const EditType t = ops[i].type;
            for (;;) {
                sp += (t == EditType::Replace || t == EditType::Delete) ? 1 : 0;
                dp += (t == EditType::Replace || t == EditType::Insert) ? 1 : 0;
                ++i;
                if (i >= ops.size() || ops[i].type != t || ops[i].src_pos != sp || ops[i].dest_pos != dp) break;
            }
            push_back(Opcode(t, sb, sp, db, dp));
        }
        if (sp < get_src_len() || dp < get_dest_len())
            push_back(Opcode(EditType::None, sp, get_src_len(), dp, get_dest_len()));
    }
    Editops as_editops() const { return Editops(*this); }
    // insertions and deletions exchanged, source and destination ranges swapped
    Opcodes inverse() const
    {
        Opcodes out = *this;
        out.set_src_len(get_dest_len());
        out.set_dest_len(get_src_len());
        for (Opcode& op : out) {
            std::swap(op.src_begin, op.dest_begin);
            std::swap(op.src_end, op.dest_end);
            if (op.type == EditType::Delete)
                op.type = EditType::Insert;
            else if (op.type == EditType::Insert)
                op.type = EditType::Delete;
        }
        return out;
    }
};

// Opcodes -> Editops: every unit of a run is one operation (None runs produce none)
inline Editops::Editops(const Opcodes& ops)
{
    set_src_len(ops.get_src_len());
    set_dest_len(ops.get_dest_len());
    for (const Opcode& op : ops) {
        switch (op.type) {
        case EditType::None: break;
        case EditType::Replace:
            for (size_t k = 0; k < op.src_end - op.src_begin; ++k)
                push_back(EditOp(EditType::Replace, op.src_begin + k, op.dest_begin + k));
            break;
        case EditType::Insert:
            for (size_t k = 0; k < op.dest_end - op.dest_begin; ++k)
                push_back(EditOp(EditType::Insert, op.src_begin, op.dest_begin + k));
            break;
        case EditType::Delete:
            for (size_t k = 0; k < op.src_end - op.src_begin; ++k)
                push_back(EditOp(EditType::Delete, op.src_begin + k, op.dest_begin));
            break;
        }
    }
}

inline Opcodes as_opcodes(const Editops& ops) { return Opcodes(ops); }

// editops_apply / opcodes_apply (details/types.hpp): s1 with the operations applied,
// taking inserted and replacing units from s2
template <typename CharT, typename Sentence1, typename Sentence2>
std::basic_string<CharT> editops_apply_str(const Editops& ops, const Sentence1& s1, const Sentence2& s2)
{
    std::basic_string<CharT> out;
    size_t src = 0;
    const auto b1 = std::begin(s1);
    const auto b2 = std::begin(s2);
    const size_t n1 = static_cast<size_t>(std::distance(std::begin(s1), std::end(s1)));
    for (const EditOp& op : ops) {
        for (; src < op.src_pos; ++src) out.push_back(static_cast<CharT>(b1[src]));
        switch (op.type) {
        case EditType::None: break;
        case EditType::Replace:
            out.push_back(static_cast<CharT>(b2[op.dest_pos]));
            ++src;
            break;
        case EditType::Insert: out.push_back(static_cast<CharT>(b2[op.dest_pos])); break;
        case EditType::Delete: ++src; break;
        }
    }
    for (; src < n1; ++src) out.push_back(static_cast<CharT>(b1[src]));
    return out;
}

template <typename CharT, typename Sentence1, typename Sentence2>
std::basic_string<CharT> opcodes_apply_str(const Opcodes& ops, const Sentence1& s1, const Sentence2& s2)
{
    std::basic_string<CharT> out;
    const auto b1 = std::begin(s1);
    const auto b2 = std::begin(s2);
    for (const Opcode& op : ops) {
        if (op.type == EditType::None)
            for (size_t k = op.src_begin; k < op.src_end; ++k) out.push_back(static_cast<CharT>(b1[k]));
        else if (op.type == EditType::Replace || op.type == EditType::Insert)
            for (size_t k = op.dest_begin; k < op.dest_end; ++k) out.push_back(static_cast<CharT>(b2[k]));
    }
    return out;
}

namespace detail {

// One pair through rfa_editops.
inline Editops editops_of(int metric, Packed a, Packed b)
{
    const int w = std::max(a.width, b.width);
    rfa_context* ctx = context();
    DeviceStrings da(ctx, a, w), db(ctx, b, w);
    uint64_t off[2] = {0, 0};
    // a script never holds more operations than the two strings have units
    std::vector<rfa_editop> raw(std::max<uint64_t>(1, a.off[1] + b.off[1]));
    check(rfa_editops(ctx, da.get(), db.get(), metric, off, raw.data(), raw.size()));
    Editops ops;
    ops.reserve(off[1]);
    for (uint64_t i = 0; i < off[1]; ++i)
        ops.push_back(EditOp(static_cast<EditType>(raw[i].type), raw[i].src_pos, raw[i].dest_pos));
    ops.set_src_len(a.off[1]);
    ops.set_dest_len(b.off[1]);
    return ops;
}

template <typename It1, typename It2>
Editops editops_of(int metric, It1 first1, It1 last1, It2 first2, It2 last2)
{
    Packed a, b;
    a.add(first1, last1);
    b.add(first2, last2);
    return editops_of(metric, std::move(a), std::move(b));
}

} // namespace detail

// levenshtein_editops (Levenshtein.hpp:282-295): pairs whose first string has at most 4096
// code units once the common prefix and suffix are removed, and that the reference does
// not split by Hirschberg (|s1| * |s2| < 2^22 or |s2| < 10 past 64 units); score_hint
// does not change the result
template <typename InputIt1, typename InputIt2>
Editops levenshtein_editops(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                            size_t score_hint = std::numeric_limits<size_t>::max())
{
    (void)score_hint;
    return detail::editops_of(RFA_LEVENSHTEIN, first1, last1, first2, last2);
}
template <typename Sentence1, typename Sentence2>
Editops levenshtein_editops(const Sentence1& s1, const Sentence2& s2,
                            size_t score_hint = std::numeric_limits<size_t>::max())
{
    return levenshtein_editops(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2),
                               score_hint);
}

// hamming_editops (Hamming.hpp:77-90): with pad = false, strings of different lengths
// throw std::invalid_argument; score_hint does not change the result
template <typename InputIt1, typename InputIt2>
Editops hamming_editops(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2, bool pad_ = true,
                        size_t score_hint = std::numeric_limits<size_t>::max())
{
    (void)score_hint;
    if (!pad_ && std::distance(first1, last1) != std::distance(first2, last2))
        throw std::invalid_argument("Sequences are not the same length.");
    return detail::editops_of(RFA_HAMMING, first1, last1, first2, last2);
}
template <typename Sentence1, typename Sentence2>
Editops hamming_editops(const Sentence1& s1, const Sentence2& s2, bool pad_ = true,
                        size_t score_hint = std::numeric_limits<size_t>::max())
{
    return hamming_editops(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2), pad_,
                           score_hint);
}

// lcs_seq_editops (LCSseq.hpp:65-75) and indel_editops (Indel.hpp:65-75, the same script):
// pairs whose first string has at most 4096 code units once the common affixes are removed
template <typename InputIt1, typename InputIt2>
Editops lcs_seq_editops(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2)
{
    return detail::editops_of(RFA_LCSSEQ, first1, last1, first2, last2);
}
template <typename Sentence1, typename Sentence2>
Editops lcs_seq_editops(const Sentence1& s1, const Sentence2& s2)
{
    return lcs_seq_editops(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2));
}
template <typename InputIt1, typename InputIt2>
Editops indel_editops(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2)
{
    return lcs_seq_editops(first1, last1, first2, last2);
}
template <typename Sentence1, typename Sentence2>
Editops indel_editops(const Sentence1& s1, const Sentence2& s2)
{
    return lcs_seq_editops(s1, s2);
}

namespace fuzz {

namespace detail {
using namespace rapidfuzz_amd::detail;

// the fuzz scorers on one cached query: similarity in [0, 100] (fuzz.hpp:82-118, 169-200)
template <int Metric>
struct CachedFuzz {
    template <typename InputIt1>
    CachedFuzz(InputIt1 first1, InputIt1 last1, bool signed8 = false)
    {
        Packed q;
        q.add(first1, last1);
        set_ = std::make_shared<ScorerSet>(Metric, std::move(q), rfa_weights{1, 1, 1}, 0.1, signed8);
    }

    template <typename InputIt2>
    double similarity(InputIt2 first2, InputIt2 last2, double score_cutoff = 0.0, double score_hint = 0.0) const
    {
        (void)score_hint;
        detail::Packed p;
        p.add(first2, last2);
        const int w = std::max(set_->query_width(), p.width);
        detail::DeviceStrings s(detail::context(), p, w);
        double v = 0;
        detail::check(rfa_cdist(set_->get(w), s.get(), RFA_SIMILARITY, RFA_CACHED, 0, score_cutoff, &v, 0));
        return v;
    }
    template <typename Sentence2>
    double similarity(const Sentence2& s2, double score_cutoff = 0.0, double score_hint = 0.0) const
    {
        return similarity(detail::to_begin(s2), detail::to_end(s2), score_cutoff, score_hint);
    }

    // batch: the query against every choice
    std::vector<double> similarity_many(const Choices& choices, double score_cutoff = 0.0) const
    {
        std::vector<double> out(choices.size());
        if (!choices.size()) return out;
        const int w = std::max(set_->query_width(), choices.width());
        detail::check(
            rfa_cdist(set_->get(w), choices.at_width(w), RFA_SIMILARITY, RFA_CACHED, 0, score_cutoff, out.data(), 0));
        return out;
    }

    // the extract() loop + ordering: k best (index, score) with score >= score_cutoff
    std::vector<std::pair<int64_t, double>> extract_topk(const Choices& choices, uint32_t k,
                                                         double score_cutoff = 0.0) const
    {
        std::vector<int64_t> idx(k);
        std::vector<double> sc(k);
        std::vector<std::pair<int64_t, double>> out;
        if (!k) return out;
        const int w = std::max(set_->query_width(), choices.width());
        detail::check(rfa_topk(set_->get(w), choices.at_width(w), RFA_SIMILARITY, score_cutoff, k, idx.data(),
                               sc.data()));
        for (uint32_t j = 0; j < k && idx[j] >= 0; ++j) out.emplace_back(idx[j], sc[j]);
        return out;
    }

private:
    std::shared_ptr<ScorerSet> set_;
};

} // namespace detail

// fuzz::CachedRatio (fuzz.hpp:82-118)
template <typename CharT1>
struct CachedRatio : detail::CachedFuzz<RFA_RATIO> {
    template <typename InputIt1>
    CachedRatio(InputIt1 first1, InputIt1 last1) : detail::CachedFuzz<RFA_RATIO>(first1, last1)
    {}
    template <typename Sentence1>
    CachedRatio(const Sentence1& s1) : CachedRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedPartialRatio (fuzz.hpp:169-200): the best ratio of the shorter string
// against the windows of the longer one
template <typename CharT1>
struct CachedPartialRatio : detail::CachedFuzz<RFA_PARTIAL_RATIO> {
    template <typename InputIt1>
    CachedPartialRatio(InputIt1 first1, InputIt1 last1) : detail::CachedFuzz<RFA_PARTIAL_RATIO>(first1, last1)
    {}
    template <typename Sentence1>
    CachedPartialRatio(const Sentence1& s1) : CachedPartialRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedTokenSortRatio (fuzz.hpp:282-304): ratio of the token-sorted strings.
// Tokens of `char` strings order as signed values, like the reference's comparison.
template <typename CharT1>
struct CachedTokenSortRatio : detail::CachedFuzz<RFA_TOKEN_SORT_RATIO> {
    template <typename InputIt1>
    CachedTokenSortRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_TOKEN_SORT_RATIO>(first1, last1,
                                                    std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedTokenSortRatio(const Sentence1& s1) : CachedTokenSortRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedPartialTokenSortRatio (fuzz.hpp:342-364): partial_ratio of the token-sorted strings
template <typename CharT1>
struct CachedPartialTokenSortRatio : detail::CachedFuzz<RFA_PARTIAL_TOKEN_SORT_RATIO> {
    template <typename InputIt1>
    CachedPartialTokenSortRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_PARTIAL_TOKEN_SORT_RATIO>(first1, last1,
                                                            std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedPartialTokenSortRatio(const Sentence1& s1)
        : CachedPartialTokenSortRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedQRatio (fuzz.hpp:776-796): ratio, 0 when either string is empty
template <typename CharT1>
struct CachedQRatio : detail::CachedFuzz<RFA_QRATIO> {
    template <typename InputIt1>
    CachedQRatio(InputIt1 first1, InputIt1 last1) : detail::CachedFuzz<RFA_QRATIO>(first1, last1)
    {}
    template <typename Sentence1>
    CachedQRatio(const Sentence1& s1) : CachedQRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedTokenSetRatio (fuzz.hpp:412-434) and fuzz::CachedTokenRatio (fuzz.hpp:530-557):
// token sets; similarity / similarity_many (no extract_topk).  Queries whose deduplicated token
// join exceeds 64 code units throw.
template <typename CharT1>
struct CachedTokenSetRatio : detail::CachedFuzz<RFA_TOKEN_SET_RATIO> {
    template <typename InputIt1>
    CachedTokenSetRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_TOKEN_SET_RATIO>(first1, last1, std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedTokenSetRatio(const Sentence1& s1) : CachedTokenSetRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

template <typename CharT1>
struct CachedTokenRatio : detail::CachedFuzz<RFA_TOKEN_RATIO> {
    template <typename InputIt1>
    CachedTokenRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_TOKEN_RATIO>(first1, last1, std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedTokenRatio(const Sentence1& s1) : CachedTokenRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedPartialTokenSetRatio (fuzz.hpp:471-493), fuzz::CachedPartialTokenRatio (fuzz.hpp:595-620)
template <typename CharT1>
struct CachedPartialTokenSetRatio : detail::CachedFuzz<RFA_PARTIAL_TOKEN_SET_RATIO> {
    template <typename InputIt1>
    CachedPartialTokenSetRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_PARTIAL_TOKEN_SET_RATIO>(first1, last1,
                                                           std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedPartialTokenSetRatio(const Sentence1& s1)
        : CachedPartialTokenSetRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

template <typename CharT1>
struct CachedPartialTokenRatio : detail::CachedFuzz<RFA_PARTIAL_TOKEN_RATIO> {
    template <typename InputIt1>
    CachedPartialTokenRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_PARTIAL_TOKEN_RATIO>(first1, last1,
                                                       std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedPartialTokenRatio(const Sentence1& s1) : CachedPartialTokenRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

// fuzz::CachedWRatio (fuzz.hpp:659-682, fuzz_impl.hpp:838-876): ratio, then token_ratio or
// partial_ratio and partial_token_ratio with cutoffs raised per pair; the token ratios' query
// limit applies.
template <typename CharT1>
struct CachedWRatio : detail::CachedFuzz<RFA_WRATIO> {
    template <typename InputIt1>
    CachedWRatio(InputIt1 first1, InputIt1 last1)
        : detail::CachedFuzz<RFA_WRATIO>(first1, last1, std::is_signed<CharT1>::value && sizeof(CharT1) == 1)
    {}
    template <typename Sentence1>
    CachedWRatio(const Sentence1& s1) : CachedWRatio(detail::to_begin(s1), detail::to_end(s1))
    {}
};

template <typename Sentence1>
CachedWRatio(const Sentence1& s1) -> CachedWRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedPartialTokenSetRatio(const Sentence1& s1) -> CachedPartialTokenSetRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedPartialTokenRatio(const Sentence1& s1) -> CachedPartialTokenRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedTokenSetRatio(const Sentence1& s1) -> CachedTokenSetRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedTokenRatio(const Sentence1& s1) -> CachedTokenRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedRatio(const Sentence1& s1) -> CachedRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedPartialRatio(const Sentence1& s1) -> CachedPartialRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedTokenSortRatio(const Sentence1& s1) -> CachedTokenSortRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedPartialTokenSortRatio(const Sentence1& s1) -> CachedPartialTokenSortRatio<detail::sentence_char_t<Sentence1>>;
template <typename Sentence1>
CachedQRatio(const Sentence1& s1) -> CachedQRatio<detail::sentence_char_t<Sentence1>>;

// fuzz free functions (fuzz.hpp:44-49, 161-167, 229-235, 334-340, 404-410, 463-469, 523-528,
// 587-593, 641-657, 713-718): one pair through the matching Cached* scorer
#define RFA_FREE_FUZZ(NAME, CACHED)                                                                                 \
    template <typename InputIt1, typename InputIt2>                                                                \
    double NAME(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2, double score_cutoff = 0)        \
    {                                                                                                              \
        return CACHED<rapidfuzz_amd::detail::iter_char_t<InputIt1>>(first1, last1)                                \
            .similarity(first2, last2, score_cutoff);                                                              \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    double NAME(const Sentence1& s1, const Sentence2& s2, double score_cutoff = 0)                                 \
    {                                                                                                              \
        return NAME(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2),             \
                    score_cutoff);                                                                                 \
    }
RFA_FREE_FUZZ(ratio, CachedRatio)
RFA_FREE_FUZZ(partial_ratio, CachedPartialRatio)
RFA_FREE_FUZZ(token_sort_ratio, CachedTokenSortRatio)
RFA_FREE_FUZZ(partial_token_sort_ratio, CachedPartialTokenSortRatio)
RFA_FREE_FUZZ(token_set_ratio, CachedTokenSetRatio)
RFA_FREE_FUZZ(partial_token_set_ratio, CachedPartialTokenSetRatio)
RFA_FREE_FUZZ(token_ratio, CachedTokenRatio)
RFA_FREE_FUZZ(partial_token_ratio, CachedPartialTokenRatio)
RFA_FREE_FUZZ(QRatio, CachedQRatio)
RFA_FREE_FUZZ(WRatio, CachedWRatio)
#undef RFA_FREE_FUZZ

// partial_ratio_alignment (fuzz.hpp:128-134): the best partial ratio and the windows of both
// strings it was found at (src: the first string); the shorter string may hold at most 4096
// code units
template <typename InputIt1, typename InputIt2>
ScoreAlignment<double> partial_ratio_alignment(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,
                                               double score_cutoff = 0)
{
    rapidfuzz_amd::detail::Packed a, b;
    a.add(first1, last1);
    b.add(first2, last2);
    const int w = std::max(a.width, b.width);
    rfa_context* ctx = rapidfuzz_amd::detail::context();
    rapidfuzz_amd::detail::DeviceStrings da(ctx, a, w), db(ctx, b, w);
    rfa_score_alignment r{};
    rapidfuzz_amd::detail::check(rfa_partial_ratio_alignment(ctx, da.get(), db.get(), score_cutoff, &r));
    return ScoreAlignment<double>(r.score, r.src_start, r.src_end, r.dest_start, r.dest_end);
}

template <typename Sentence1, typename Sentence2>
ScoreAlignment<double> partial_ratio_alignment(const Sentence1& s1, const Sentence2& s2, double score_cutoff = 0)
{
    return partial_ratio_alignment(detail::to_begin(s1), detail::to_end(s1), detail::to_begin(s2), detail::to_end(s2),
                                   score_cutoff);
}

} // namespace fuzz

// ---- experimental::Multi* -------------------------------------------------

namespace experimental {
namespace detail {

// AVX2 lane counts of the reference build (Levenshtein.hpp:303-321): a 256-bit
// vector holds 256 / (8 * element bytes) strings of up to MaxLen units.
template <int MaxLen>
constexpr size_t vec_size()
{
    static_assert(MaxLen <= 64, "expected MaxLen <= 64");
    return MaxLen <= 8 ? 32 : (MaxLen <= 16 ? 16 : (MaxLen <= 32 ? 8 : 4));
}

template <int Metric, int MaxLen, typename ResType, int64_t WorstSimilarity, int64_t WorstDistance>
class MultiBase {
public:
    explicit MultiBase(size_t count, rfa_weights w = {1, 1, 1}, double pw = 0.1)
        : input_count_(count), w_(w), pw_(pw)
    {}

    size_t result_count() const { return (input_count_ + vec_size<MaxLen>() - 1) / vec_size<MaxLen>() * vec_size<MaxLen>(); }

    template <typename Sentence1>
    void insert(const Sentence1& s1_)
    {
        insert(rapidfuzz_amd::detail::to_begin(s1_), rapidfuzz_amd::detail::to_end(s1_));
    }
    template <typename InputIt1>
    void insert(InputIt1 first1, InputIt1 last1)
    {
        if (pos_ >= input_count_) throw std::invalid_argument("out of bounds insert");
        if (std::distance(first1, last1) > MaxLen) throw std::invalid_argument("string longer than MaxLen");
        q_.add(first1, last1);
        ++pos_;
        std::lock_guard<std::mutex> lk(mu_);
        set_.reset();
    }

    // scores[i] for i < result_count(); entries past the inserted strings score an empty string
    template <typename InputIt2>
    void distance(ResType* scores, size_t score_count, InputIt2 first2, InputIt2 last2,
                  ResType score_cutoff = static_cast<ResType>(WorstDistance)) const
    {
        score(scores, score_count, first2, last2, RFA_DISTANCE, score_cutoff);
    }
    template <typename Sentence2>
    void distance(ResType* scores, size_t score_count, const Sentence2& s2,
                  ResType score_cutoff = static_cast<ResType>(WorstDistance)) const
    {
        distance(scores, score_count, rapidfuzz_amd::detail::to_begin(s2), rapidfuzz_amd::detail::to_end(s2),
                 score_cutoff);
    }
    template <typename InputIt2>
    void similarity(ResType* scores, size_t score_count, InputIt2 first2, InputIt2 last2,
                    ResType score_cutoff = static_cast<ResType>(WorstSimilarity)) const
    {
        score(scores, score_count, first2, last2, RFA_SIMILARITY, score_cutoff);
    }
    template <typename Sentence2>
    void similarity(ResType* scores, size_t score_count, const Sentence2& s2,
                    ResType score_cutoff = static_cast<ResType>(WorstSimilarity)) const
    {
        similarity(scores, score_count, rapidfuzz_amd::detail::to_begin(s2), rapidfuzz_amd::detail::to_end(s2),
                   score_cutoff);
    }
    template <typename InputIt2>
    void normalized_distance(double* scores, size_t score_count, InputIt2 first2, InputIt2 last2,
                             double score_cutoff = 1.0) const
    {
        score(scores, score_count, first2, last2, RFA_NORMALIZED_DISTANCE, score_cutoff);
    }
    template <typename Sentence2>
    void normalized_distance(double* scores, size_t score_count, const Sentence2& s2, double score_cutoff = 1.0) const
    {
        normalized_distance(scores, score_count, rapidfuzz_amd::detail::to_begin(s2),
                            rapidfuzz_amd::detail::to_end(s2), score_cutoff);
    }
    template <typename InputIt2>
    void normalized_similarity(double* scores, size_t score_count, InputIt2 first2, InputIt2 last2,
                               double score_cutoff = 0.0) const
    {
        score(scores, score_count, first2, last2, RFA_NORMALIZED_SIMILARITY, score_cutoff);
    }
    template <typename Sentence2>
    void normalized_similarity(double* scores, size_t score_count, const Sentence2& s2,
                               double score_cutoff = 0.0) const
    {
        normalized_similarity(scores, score_count, rapidfuzz_amd::detail::to_begin(s2),
                              rapidfuzz_amd::detail::to_end(s2), score_cutoff);
    }

    // batch: every text of `choices`; out[t * result_count() + i]
    template <typename Out>
    void score_many(const Choices& choices, int kind, Out* out, double cutoff_d, uint64_t cutoff_u) const
    {
        const size_t rc = result_count(), nc = choices.size();
        if (!nc) return;
        const int w = std::max(query_width(), choices.width());
        std::vector<Out> tmp(rc * nc);
        rapidfuzz_amd::detail::check(
            rfa_cdist(scorer(w), choices.at_width(w), kind, RFA_MULTI, cutoff_u, cutoff_d, tmp.data(), 0));
        for (size_t i = 0; i < rc; ++i)
            for (size_t t = 0; t < nc; ++t) out[t * rc + i] = tmp[i * nc + t];
    }

    // batch with one cutoff per inserted string (rfa_cdist_cutoffs; integer results only):
    // out[t * result_count() + i] = the score of string i against text t, cut at score_cutoffs[i]
    // (padding entries past the inserted strings take the default cutoff)
    std::vector<ResType> distance_many(const Choices& choices, const std::vector<ResType>& score_cutoffs) const
    {
        return many_cutoffs(choices, RFA_DISTANCE, score_cutoffs, static_cast<ResType>(WorstDistance));
    }
    std::vector<ResType> similarity_many(const Choices& choices, const std::vector<ResType>& score_cutoffs) const
    {
        return many_cutoffs(choices, RFA_SIMILARITY, score_cutoffs, static_cast<ResType>(WorstSimilarity));
    }

protected:
    int query_width() const { return q_.width; }

    rfa_scorer* scorer(int width) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!set_) {
            rapidfuzz_amd::detail::Packed all = q_;
            while (all.count() < result_count()) all.off.push_back(all.units.size()); // empty padding lanes
            set_ = std::make_shared<rapidfuzz_amd::detail::ScorerSet>(Metric, std::move(all), w_, pw_);
        }
        return set_->get(width);
    }

    template <typename InputIt2, typename Out, typename Cut>
    void score(Out* scores, size_t score_count, InputIt2 first2, InputIt2 last2, int kind, Cut cutoff) const
    {
        if (score_count < result_count())
            throw std::invalid_argument("scores has to have >= result_count() elements");
        rapidfuzz_amd::detail::Packed p;
        p.add(first2, last2);
        const int w = std::max(query_width(), p.width);
        rapidfuzz_amd::detail::DeviceStrings s(rapidfuzz_amd::detail::context(), p, w);
        const bool int_out = std::is_integral<Out>::value;
        std::vector<uint64_t> tmp(result_count());
        rapidfuzz_amd::detail::check(rfa_cdist(scorer(w), s.get(), kind, RFA_MULTI,
                                               int_out ? static_cast<uint64_t>(cutoff) : 0,
                                               int_out ? 0.0 : static_cast<double>(cutoff), tmp.data(), 0));
        for (size_t i = 0; i < result_count(); ++i) {
            if constexpr (std::is_integral<Out>::value)
                scores[i] = static_cast<Out>(tmp[i]);
            else
                std::memcpy(&scores[i], &tmp[i], sizeof(double));
        }
    }

    std::vector<ResType> many_cutoffs(const Choices& choices, int kind, const std::vector<ResType>& cuts,
                                      ResType pad_cut) const
    {
        static_assert(std::is_integral<ResType>::value, "per-string cutoffs apply to integer results");
        if (cuts.size() != pos_) throw std::invalid_argument("score_cutoffs needs one entry per inserted string");
        const size_t rc = result_count(), nc = choices.size();
        std::vector<ResType> out(rc * nc);
        if (!nc) return out;
        std::vector<uint64_t> cu(rc, static_cast<uint64_t>(pad_cut));
        for (size_t i = 0; i < cuts.size(); ++i) cu[i] = static_cast<uint64_t>(cuts[i]);
        const int w = std::max(query_width(), choices.width());
        std::vector<uint64_t> tmp(rc * nc);
        rapidfuzz_amd::detail::check(
            rfa_cdist_cutoffs(scorer(w), choices.at_width(w), kind, RFA_MULTI, cu.data(), tmp.data(), 0));
        for (size_t i = 0; i < rc; ++i)
            for (size_t t = 0; t < nc; ++t) out[t * rc + i] = static_cast<ResType>(tmp[i * nc + t]);
        return out;
    }

    size_t input_count_;
    size_t pos_ = 0;
    rapidfuzz_amd::detail::Packed q_;
    rfa_weights w_;
    double pw_;
    mutable std::mutex mu_;
    mutable std::shared_ptr<rapidfuzz_amd::detail::ScorerSet> set_;
};

constexpr int64_t kIntMax = std::numeric_limits<int64_t>::max();

} // namespace detail

template <int MaxLen>
struct MultiLevenshtein : detail::MultiBase<RFA_LEVENSHTEIN, MaxLen, size_t, 0, detail::kIntMax> {
    MultiLevenshtein(size_t count, LevenshteinWeightTable aWeights = {1, 1, 1})
        : detail::MultiBase<RFA_LEVENSHTEIN, MaxLen, size_t, 0, detail::kIntMax>(
              count, {aWeights.insert_cost, aWeights.delete_cost, aWeights.replace_cost})
    {
        // Levenshtein.hpp:334-340
        if (aWeights.delete_cost != 1 || aWeights.insert_cost != 1 || aWeights.replace_cost > 2)
            throw std::invalid_argument("unsupported weights");
    }
};

template <int MaxLen>
struct MultiIndel : detail::MultiBase<RFA_INDEL, MaxLen, size_t, 0, detail::kIntMax> {
    explicit MultiIndel(size_t count) : detail::MultiBase<RFA_INDEL, MaxLen, size_t, 0, detail::kIntMax>(count) {}
};

template <int MaxLen>
struct MultiLCSseq : detail::MultiBase<RFA_LCSSEQ, MaxLen, size_t, 0, detail::kIntMax> {
    explicit MultiLCSseq(size_t count) : detail::MultiBase<RFA_LCSSEQ, MaxLen, size_t, 0, detail::kIntMax>(count) {}
};

template <int MaxLen>
struct MultiOSA : detail::MultiBase<RFA_OSA, MaxLen, size_t, 0, detail::kIntMax> {
    explicit MultiOSA(size_t count) : detail::MultiBase<RFA_OSA, MaxLen, size_t, 0, detail::kIntMax>(count) {}
};

template <int MaxLen>
struct MultiJaro : detail::MultiBase<RFA_JARO, MaxLen, double, 0, 1> {
    explicit MultiJaro(size_t count) : detail::MultiBase<RFA_JARO, MaxLen, double, 0, 1>(count) {}
};

template <int MaxLen>
struct MultiJaroWinkler : detail::MultiBase<RFA_JARO_WINKLER, MaxLen, double, 0, 1> {
    explicit MultiJaroWinkler(size_t count, double prefix_weight = 0.1)
        : detail::MultiBase<RFA_JARO_WINKLER, MaxLen, double, 0, 1>(count, {1, 1, 1}, prefix_weight)
    {}
};

// experimental::CachedDamerauLevenshtein (DamerauLevenshtein.hpp:112-140)
template <typename CharT1>
struct CachedDamerauLevenshtein
    : rapidfuzz_amd::detail::CachedBase<RFA_DAMERAU_LEVENSHTEIN, size_t, 0, rapidfuzz_amd::detail::kIntMax> {
    template <typename Sentence1>
    explicit CachedDamerauLevenshtein(const Sentence1& s1_)
        : CachedDamerauLevenshtein(rapidfuzz_amd::detail::to_begin(s1_), rapidfuzz_amd::detail::to_end(s1_))
    {}
    template <typename InputIt1>
    CachedDamerauLevenshtein(InputIt1 first1, InputIt1 last1)
    {
        this->init(first1, last1, {1, 1, 1}, 0.1);
    }
};

template <typename Sentence1>
explicit CachedDamerauLevenshtein(const Sentence1& s1_)
    -> CachedDamerauLevenshtein<rapidfuzz_amd::detail::sentence_char_t<Sentence1>>;

// damerau_levenshtein_* (DamerauLevenshtein.hpp:31-112): one pair through a one-query scorer
#define RFA_FREE_DL(KIND, RES, CUT)                                                                                 \
    template <typename InputIt1, typename InputIt2>                                                                \
    RES damerau_levenshtein_##KIND(InputIt1 first1, InputIt1 last1, InputIt2 first2, InputIt2 last2,             \
                                   RES score_cutoff = CUT)                                                         \
    {                                                                                                              \
        return CachedDamerauLevenshtein<rapidfuzz_amd::detail::iter_char_t<InputIt1>>(first1, last1)              \
            .KIND(first2, last2, score_cutoff);                                                                    \
    }                                                                                                              \
    template <typename Sentence1, typename Sentence2>                                                              \
    RES damerau_levenshtein_##KIND(const Sentence1& s1, const Sentence2& s2, RES score_cutoff = CUT)               \
    {                                                                                                              \
        return damerau_levenshtein_##KIND(rapidfuzz_amd::detail::to_begin(s1), rapidfuzz_amd::detail::to_end(s1),  \
                                          rapidfuzz_amd::detail::to_begin(s2), rapidfuzz_amd::detail::to_end(s2),  \
                                          score_cutoff);                                                           \
    }
RFA_FREE_DL(distance, size_t, std::numeric_limits<size_t>::max())
RFA_FREE_DL(similarity, size_t, 0)
RFA_FREE_DL(normalized_distance, double, 1.0)
RFA_FREE_DL(normalized_similarity, double, 0.0)
#undef RFA_FREE_DL

} // namespace experimental

namespace fuzz {
namespace experimental {

// fuzz.hpp:43-93: similarity() only, on the 0-100 scale
template <int MaxLen>
struct MultiRatio : rapidfuzz_amd::experimental::detail::MultiBase<RFA_RATIO, MaxLen, double, 0, 100> {
    explicit MultiRatio(size_t count)
        : rapidfuzz_amd::experimental::detail::MultiBase<RFA_RATIO, MaxLen, double, 0, 100>(count)
    {}
};

} // namespace experimental
} // namespace fuzz

// ---- independent pairs in one launch (rfa_pairs) --------------------------------
// out[i] = <metrics[i]>_<kind>(a[i], b[i], score_cutoff) for the free functions
// levenshtein_* (weights {1, 1, 1}), indel_*, lcs_seq_* and osa_*: the batched form of
// per-pair calls (Levenshtein.hpp:140-265, Indel.hpp:12-60, LCSseq.hpp:12-60, OSA.hpp:36-108),
// pairs of any lengths mixed in one call.
enum class PairMetric : int32_t {
    Levenshtein = RFA_LEVENSHTEIN,
    Indel = RFA_INDEL,
    LCSseq = RFA_LCSSEQ,
    OSA = RFA_OSA
};

namespace detail {
template <typename Out>
std::vector<Out> run_pairs(const Choices& a, const Choices& b, const std::vector<int32_t>& metrics, int kind,
                           uint64_t cu, double cd, int schedule)
{
    if (a.size() != b.size()) throw std::invalid_argument("a and b must hold the same number of strings");
    if (metrics.size() != a.size()) throw std::invalid_argument("metrics needs one entry per pair");
    std::vector<Out> out(a.size());
    if (!a.size()) return out;
    const int w = std::max(a.width(), b.width());
    check(rfa_pairs(context(), a.at_width(w), b.at_width(w), metrics.data(), kind, cu, cd, schedule, out.data(), 0));
    return out;
}
inline std::vector<int32_t> same_metric(size_t n, PairMetric m) { return std::vector<int32_t>(n, static_cast<int32_t>(m)); }
} // namespace detail

inline std::vector<size_t> pairs_distance(const Choices& a, const Choices& b, const std::vector<int32_t>& metrics,
                                          size_t score_cutoff = std::numeric_limits<size_t>::max(),
                                          int schedule = RFA_SCHEDULE_BUCKETED)
{
    const uint64_t cu = std::min<uint64_t>(score_cutoff, static_cast<uint64_t>(detail::kIntMax));
    auto v = detail::run_pairs<uint64_t>(a, b, metrics, RFA_DISTANCE, cu, 0.0, schedule);
    return std::vector<size_t>(v.begin(), v.end());
}
inline std::vector<size_t> pairs_similarity(const Choices& a, const Choices& b, const std::vector<int32_t>& metrics,
                                            size_t score_cutoff = 0, int schedule = RFA_SCHEDULE_BUCKETED)
{
    auto v = detail::run_pairs<uint64_t>(a, b, metrics, RFA_SIMILARITY, score_cutoff, 0.0, schedule);
    return std::vector<size_t>(v.begin(), v.end());
}
inline std::vector<double> pairs_normalized_distance(const Choices& a, const Choices& b,
                                                     const std::vector<int32_t>& metrics, double score_cutoff = 1.0,
                                                     int schedule = RFA_SCHEDULE_BUCKETED)
{
    return detail::run_pairs<double>(a, b, metrics, RFA_NORMALIZED_DISTANCE, 0, score_cutoff, schedule);
}
inline std::vector<double> pairs_normalized_similarity(const Choices& a, const Choices& b,
                                                       const std::vector<int32_t>& metrics,
                                                       double score_cutoff = 0.0,
                                                       int schedule = RFA_SCHEDULE_BUCKETED)
{
    return detail::run_pairs<double>(a, b, metrics, RFA_NORMALIZED_SIMILARITY, 0, score_cutoff, schedule);
}
// one metric for every pair
inline std::vector<size_t> pairs_distance(const Choices& a, const Choices& b, PairMetric m,
                                          size_t score_cutoff = std::numeric_limits<size_t>::max())
{
    return pairs_distance(a, b, detail::same_metric(a.size(), m), score_cutoff);
}
inline std::vector<size_t> pairs_similarity(const Choices& a, const Choices& b, PairMetric m, size_t score_cutoff = 0)
{
    return pairs_similarity(a, b, detail::same_metric(a.size(), m), score_cutoff);
}
inline std::vector<double> pairs_normalized_distance(const Choices& a, const Choices& b, PairMetric m,
                                                     double score_cutoff = 1.0)
{
    return pairs_normalized_distance(a, b, detail::same_metric(a.size(), m), score_cutoff);
}
inline std::vector<double> pairs_normalized_similarity(const Choices& a, const Choices& b, PairMetric m,
                                                       double score_cutoff = 0.0)
{
    return pairs_normalized_similarity(a, b, detail::same_metric(a.size(), m), score_cutoff);
}

// ---- request server (rfa_server_*) -------------------------------------------------
// A device-resident choice set answering top-k requests with continuous batching: the
// README's extract() loop (README.md:153-214) as a service.  No reference class counterpart.
struct ServerConfig {
    int metric = RFA_RATIO;
    int kind = RFA_SIMILARITY; // a double similarity: normalized_similarity, or similarity of Jaro / JW / ratio
    double score_cutoff = 0.0;
    uint32_t k = 5;
    uint32_t max_batch = 64;
    uint32_t slots = 4;
    uint32_t candidate_cap = 65536;
    double prefix_weight = 0.1;
    LevenshteinWeightTable weights = {1, 1, 1};
};

class Server {
public:
    explicit Server(const Choices& choices, const ServerConfig& cfg = ServerConfig())
        : width_(choices.width()), k_(cfg.k)
    {
        rfa_server_config c{};
        c.metric = cfg.metric;
        c.kind = cfg.kind;
        c.score_cutoff = cfg.score_cutoff;
        c.k = cfg.k;
        c.max_batch = cfg.max_batch;
        c.slots = cfg.slots;
        c.candidate_cap = cfg.candidate_cap;
        c.prefix_weight = cfg.prefix_weight;
        c.weights = {cfg.weights.insert_cost, cfg.weights.delete_cost, cfg.weights.replace_cost};
        detail::check(rfa_server_create(detail::context(), choices.at_width(width_), &c, &h_));
    }
    ~Server() { rfa_server_destroy(h_); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // queues one query (code units of at most the choice set's width); returns a ticket at once
    template <typename InputIt>
    uint64_t submit(InputIt first, InputIt last) const
    {
        detail::Packed p;
        p.add(first, last);
        if (p.width > width_) throw std::invalid_argument("query code units are wider than the choice set's");
        std::vector<uint8_t> bytes(std::max<size_t>(p.units.size() * static_cast<size_t>(width_), 1));
        for (size_t i = 0; i < p.units.size(); ++i) {
            const uint32_t u = p.units[i];
            if (width_ == 1)
                bytes[i] = static_cast<uint8_t>(u);
            else if (width_ == 2) {
                const uint16_t v = static_cast<uint16_t>(u);
                std::memcpy(&bytes[i * 2], &v, 2);
            }
            else
                std::memcpy(&bytes[i * 4], &u, 4);
        }
        uint64_t t = 0;
        detail::check(rfa_server_submit(h_, bytes.data(), p.units.size(), &t));
        return t;
    }
    template <typename Sentence>
    uint64_t submit(const Sentence& q) const
    {
        return submit(detail::to_begin(q), detail::to_end(q));
    }

    // blocks until the request is answered: (choice index, score), best first, at most k
    std::vector<std::pair<int64_t, double>> wait(uint64_t ticket) const
    {
        std::vector<int64_t> idx(std::max<uint32_t>(k_, 1));
        std::vector<double> sc(idx.size());
        detail::check(rfa_server_wait(h_, ticket, idx.data(), sc.data()));
        std::vector<std::pair<int64_t, double>> out;
        for (uint32_t j = 0; j < k_ && idx[j] >= 0; ++j) out.emplace_back(idx[j], sc[j]);
        return out;
    }
    template <typename Sentence>
    std::vector<std::pair<int64_t, double>> query(const Sentence& q) const
    {
        return wait(submit(q));
    }

    rfa_server_stats stats() const
    {
        rfa_server_stats st{};
        detail::check(rfa_server_get_stats(h_, &st));
        return st;
    }

private:
    int width_;
    uint32_t k_;
    rfa_server* h_ = nullptr;
};

} // namespace rapidfuzz_amd
