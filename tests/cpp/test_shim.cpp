// test_shim.cpp -- the C++ drop-in classes (rapidfuzz_amd.hpp) on the GPU.
//
// Known answers from the reference's own tests (tests-Levenshtein.cpp:144-179,
// tests-Indel.cpp:204-215, tests-OSA.cpp:62-97, tests-JaroWinkler.cpp:181-182,
// tests-fuzz.cpp:63-76, tests-Hamming.cpp:99-108), batch results equal to one-pair results, Multi*
// padding and error checks, and extract_topk.  Exit status: 0 pass, 1 fail,
// 77 no GPU device.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "rapidfuzz_amd/rapidfuzz_amd.hpp"

namespace rfa = rapidfuzz_amd;

static int g_fail = 0, g_checks = 0;

#define CHECK(cond)                                                                                                \
    do {                                                                                                           \
        ++g_checks;                                                                                                \
        if (!(cond)) {                                                                                             \
            ++g_fail;                                                                                              \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                                            \
        }                                                                                                          \
    } while (0)

static bool near(double a, double b) { return std::fabs(a - b) < 1e-6; }

template <typename F>
static bool throws_invalid(F&& f)
{
    try {
        f();
    }
    catch (const std::invalid_argument&) {
        return true;
    }
    return false;
}

// editops_apply_str (details/types.hpp) restated for the checks below
static std::string apply_editops(const rfa::Editops& ops, const std::string& s1, const std::string& s2)
{
    std::string out;
    size_t src = 0;
    for (size_t k = 0; k < ops.size(); ++k) {
        while (src < ops[k].src_pos) out.push_back(s1[src++]);
        if (ops[k].type == rfa::EditType::Delete)
            ++src;
        else if (ops[k].type == rfa::EditType::Insert)
            out.push_back(s2[ops[k].dest_pos]);
        else if (ops[k].type == rfa::EditType::Replace) {
            out.push_back(s2[ops[k].dest_pos]);
            ++src;
        }
    }
    while (src < s1.size()) out.push_back(s1[src++]);
    return out;
}

static std::string random_string(std::mt19937& rng, size_t lo, size_t hi)
{
    std::uniform_int_distribution<size_t> len(lo, hi);
    std::uniform_int_distribution<int> ch('a', 'e');
    std::string s(len(rng), 'a');
    for (auto& c : s) c = static_cast<char>(ch(rng));
    return s;
}

int main()
{
    try {
        // ---- known answers ----
        {
            rfa::CachedLevenshtein<char> lev(std::string("South Korea"));
            CHECK(lev.distance(std::string("North Korea")) == 2);
            CHECK(lev.distance(std::string("North Korea"), 1) == 2);
            CHECK(lev.distance(std::string("North Korea"), 0) == 1);
            rfa::CachedLevenshtein<char> lev112(std::string("South Korea"), rfa::LevenshteinWeightTable{1, 1, 2});
            CHECK(lev112.distance(std::string("North Korea")) == 4);
            CHECK(lev112.distance(std::string("North Korea"), 2) == 3);
            rfa::CachedLevenshtein<char> aabc("aabc");
            CHECK(aabc.distance("cccd") == 4);
            CHECK(aabc.distance("cccd", 2) == 3);
            // general weights {insert 1, delete 2, replace 3}: two replacements and an insertion,
            // out of a maximum of 6 * 2 + 7 * 1 = 19
            rfa::CachedLevenshtein<char> w123(std::string("kitten"), rfa::LevenshteinWeightTable{1, 2, 3});
            CHECK(w123.distance(std::string("sitting")) == 7);
            CHECK(w123.distance(std::string("sitting"), 5) == 6);
            CHECK(w123.distance(std::string("kitchen")) == 4);
            CHECK(std::fabs(w123.normalized_similarity(std::string("sitting")) - (1.0 - 7.0 / 19.0)) < 1e-12);
        }
        {
            // tests-Levenshtein.cpp:235-244, 294-309 and tests-Indel.cpp:275-285 (apply-based)
            const std::string s = "Lorem ipsum.", d = "XYZLorem ABC iPsum";
            rfa::Editops ops = rfa::levenshtein_editops(s, d);
            CHECK(apply_editops(ops, s, d) == d);
            CHECK(ops.get_src_len() == s.size());
            CHECK(ops.get_dest_len() == d.size());
            CHECK(ops.size() == rfa::CachedLevenshtein<char>(s).distance(d));
            const std::string b1 = "b", b2 = "aaaaaaaaaaaaaaaabbaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa";
            CHECK(apply_editops(rfa::levenshtein_editops(b1, b2), b1, b2) == b2);
            CHECK(apply_editops(rfa::levenshtein_editops(std::string("aa"), std::string("abb")), "aa", "abb") == "abb");
            rfa::Editops ind = rfa::indel_editops(s, d);
            CHECK(apply_editops(ind, s, d) == d);
            CHECK(ind.size() == rfa::CachedIndel<char>(s).distance(d));
            CHECK(ind == rfa::lcs_seq_editops(s, d));
            // Opcodes (details/types.hpp:357-565): round trip and apply
            rfa::Opcodes oc(ops);
            CHECK(oc.as_editops() == ops);
            CHECK(rfa::opcodes_apply_str<char>(oc, s, d) == d);
            CHECK(rfa::editops_apply_str<char>(ops, s, d) == d);
            CHECK(oc.get_src_len() == s.size() && oc.get_dest_len() == d.size());
            // tests-Hamming.cpp:112-127 (pad = true)
            rfa::Editops hops = rfa::hamming_editops(s, d);
            CHECK(rfa::editops_apply_str<char>(hops, s, d) == d);
            CHECK(hops.get_src_len() == s.size() && hops.get_dest_len() == d.size());
            CHECK(rfa::editops_apply_str<char>(rfa::hamming_editops(d, s), d, s) == s);
            CHECK(hops.size() == rfa::CachedHamming<char>(s).distance(d));
            CHECK(throws_invalid([&] { rfa::hamming_editops(s, d, false); }));
        }
        {
            // tests-fuzz.cpp:187-200 (issue 231) and the swap of the arguments
            const std::string s1 = "er merkantilismus f/rderte handel und verkehr mit teils marktkonformen, teils "
                                   "dirigistischen ma_nahmen.";
            const std::string s2 = "ils marktkonformen, teils dirigistischen ma_nahmen. an der schwelle zum 19. "
                                   "jahrhundert entstand ein neu";
            const auto al = rfa::fuzz::partial_ratio_alignment(s1, s2);
            CHECK(std::fabs(al.score - 66.2337662) < 1e-6);
            CHECK(al.src_start == 0 && al.src_end == 103 && al.dest_start == 0 && al.dest_end == 51);
            const auto sw = rfa::fuzz::partial_ratio_alignment(s2, s1);
            CHECK(sw == rfa::ScoreAlignment<double>(al.score, 0, 51, 0, 103));
            CHECK(al.score == rfa::fuzz::partial_ratio(s1, s2));
            CHECK(rfa::fuzz::partial_ratio_alignment(std::string(""), std::string("")) ==
                  rfa::ScoreAlignment<double>(100.0, 0, 0, 0, 0));
            CHECK(rfa::fuzz::partial_ratio_alignment(s1, s2, 101.0) == rfa::ScoreAlignment<double>(0.0, 0, 103, 0, 103));
        }
        {
            // tests-DamerauLevenshtein.cpp:106-130
            rfa::experimental::CachedDamerauLevenshtein dl(std::string("abaa"));
            CHECK(dl.distance(std::string("baaa")) == 1);
            CHECK(std::fabs(dl.normalized_similarity(std::string("baaa")) - 0.75) < 1e-12);
            rfa::experimental::CachedDamerauLevenshtein<char> ca(std::string("CA"));
            CHECK(ca.distance(std::string("ABC")) == 2);
            CHECK(ca.distance(std::string("ABC"), 1) == 2);
            CHECK(std::fabs(ca.normalized_similarity(std::string("ABC")) - 1.0 / 3.0) < 1e-12);
        }
        {
            rfa::CachedIndel<char> indel(std::string("aaaa"));
            CHECK(indel.distance(std::string("aaaa")) == 0);
            CHECK(indel.similarity(std::string("aaaa")) == 8);
            CHECK(indel.distance(std::string("bbbb")) == 8);
            CHECK(indel.normalized_similarity(std::string("aaaa")) == 1.0);
        }
        {
            rfa::CachedOSA<char> osa(std::string("CA"));
            CHECK(osa.distance(std::string("ABC")) == 3);
            CHECK(osa.distance(std::string("AC")) == 1);
            rfa::CachedOSA<char> aaaa(std::string("aaaa"));
            CHECK(aaaa.distance(std::string("")) == 4);
            CHECK(aaaa.distance(std::string(""), 1) == 2);
        }
        {
            rfa::CachedJaroWinkler<char> jw(std::string("0"));
            CHECK(near(jw.similarity(std::string("00")), 0.85));
            rfa::CachedJaroWinkler<char> jw2(std::string("00"));
            CHECK(near(jw2.similarity(std::string("00")), 1.0));
        }
        {
            rfa::fuzz::CachedRatio<char> ratio(std::string("new york mets"));
            CHECK(near(ratio.similarity(std::string("new york mets")), 100.0));
            CHECK(near(ratio.similarity(std::string("the wonderful new york mets")), 65.0));
            // tests-fuzz.cpp:77-79, 157-158
            rfa::fuzz::CachedPartialRatio<char> partial(std::string("new york mets"));
            CHECK(near(partial.similarity(std::string("new york mets")), 100.0));
            CHECK(near(partial.similarity(std::string("the wonderful new york mets")), 100.0));
            rfa::fuzz::CachedPartialRatio p001(std::string("001"));
            CHECK(near(p001.similarity(std::string("220222")), 33.3333333));
            CHECK(near(rfa::fuzz::CachedPartialRatio<char>(std::string("physics 2 vid"))
                           .similarity(std::string("study physics physics 2 video")),
                       100.0));
            // tests-fuzz.cpp:84-85, 125-151
            rfa::fuzz::CachedTokenSortRatio<char> ts(std::string("new york mets"));
            CHECK(near(ts.similarity(std::string("new york mets")), 100.0));
            CHECK(near(ts.similarity(std::string("mets york new")), 100.0));
            CHECK(near(rfa::fuzz::CachedTokenSortRatio<char>(std::string("")).similarity(std::string("")), 100.0));
            CHECK(near(rfa::fuzz::CachedPartialTokenSortRatio<char>(std::string("")).similarity(std::string("")),
                       100.0));
            CHECK(near(rfa::fuzz::CachedQRatio<char>(std::string("")).similarity(std::string("")), 0.0));
            CHECK(near(rfa::fuzz::CachedQRatio<char>(std::string("test")).similarity(std::string("")), 0.0));
            // `char` tokens order as signed bytes: "\xe9x" sorts before "ab", so the join is
            // "\xe9x ab" (88.9 against "\xe9xab"; unsigned order would give "ab \xe9x", 44.4)
            // tests-fuzz.cpp:88-94, 125-135
            const std::string s4 = "new york mets vs atlanta braves", s5 = "atlanta braves vs new york mets";
            CHECK(near(rfa::fuzz::CachedTokenSetRatio<char>(s4).similarity(s5), 100.0));
            CHECK(near(rfa::fuzz::CachedTokenSetRatio<char>(std::string("a{")).similarity(std::string("{b")), 50.0));
            CHECK(near(rfa::fuzz::CachedTokenSetRatio<char>(std::string("")).similarity(std::string("")), 0.0));
            CHECK(near(rfa::fuzz::CachedTokenRatio<char>(std::string("")).similarity(std::string("")), 100.0));
            CHECK(near(rfa::fuzz::CachedTokenRatio<char>(s4).similarity(s5), 100.0));
            // tests-fuzz.cpp:96-99, 125-135
            CHECK(near(rfa::fuzz::CachedPartialTokenSetRatio<char>(s4).similarity(
                           std::string("new york city mets - atlanta braves")),
                       100.0));
            CHECK(near(rfa::fuzz::CachedPartialTokenSetRatio<char>(std::string("")).similarity(std::string("")), 0.0));
            CHECK(near(rfa::fuzz::CachedPartialTokenRatio<char>(std::string("")).similarity(std::string("")), 100.0));
            rfa::fuzz::CachedTokenSortRatio<char> acc(std::string("\xe9x ab"));
            CHECK(near(acc.similarity(std::string("\xe9xab")), 88.8888889));
            rfa::fuzz::CachedTokenSortRatio<unsigned char> accu(std::basic_string<unsigned char>(
                reinterpret_cast<const unsigned char*>("\xe9x ab")));
            CHECK(near(accu.similarity(std::basic_string<unsigned char>(
                           reinterpret_cast<const unsigned char*>("\xe9xab"))),
                       44.4444444));
        }
        {
            // deduction guides and wide code units against narrow ones
            rfa::CachedLevenshtein lev(std::u32string(U"käse"));
            CHECK(lev.distance(std::u32string(U"kase")) == 1);
            CHECK(lev.distance(std::string("k\xe4se")) == 0); // 0xE4 as an unsigned 8-bit unit
        }

        // ---- batch == one pair at a time ----
        std::mt19937 rng(7);
        std::vector<std::string> choices;
        for (int i = 0; i < 300; ++i) choices.push_back(random_string(rng, 0, 90));
        const rfa::Choices set(choices);
        for (int qi = 0; qi < 4; ++qi) {
            const std::string q = random_string(rng, 0, 80);
            rfa::CachedLevenshtein<char> lev(q);
            rfa::CachedIndel<char> indel(q);
            rfa::CachedOSA<char> osa(q);
            rfa::CachedJaroWinkler<char> jw(q);
            rfa::fuzz::CachedRatio<char> ratio(q);
            rfa::fuzz::CachedPartialRatio<char> partial(q);
            const auto d = lev.distance_many(set);
            const auto dc = lev.distance_many(set, 5);
            const auto ns = indel.normalized_similarity_many(set, 0.6);
            const auto od = osa.distance_many(set);
            const auto js = jw.similarity_many(set);
            const auto rs = ratio.similarity_many(set, 50.0);
            const auto ps = partial.similarity_many(set, 50.0);
            for (size_t c = 0; c < choices.size(); c += 7) {
                CHECK(d[c] == lev.distance(choices[c]));
                CHECK(dc[c] == lev.distance(choices[c], 5));
                CHECK(ns[c] == indel.normalized_similarity(choices[c], 0.6));
                CHECK(od[c] == osa.distance(choices[c]));
                CHECK(js[c] == jw.similarity(choices[c]));
                CHECK(rs[c] == ratio.similarity(choices[c], 50.0));
                CHECK(ps[c] == partial.similarity(choices[c], 50.0));
            }
            // extract(): the k best by (score desc, index asc) among scores >= cutoff
            const auto top = ratio.extract_topk(set, 5, 40.0);
            std::vector<std::pair<int64_t, double>> expect;
            for (size_t c = 0; c < choices.size(); ++c) {
                const double s = ratio.similarity_many(set, 40.0)[c];
                if (s >= 40.0) expect.emplace_back((int64_t)c, s);
            }
            std::stable_sort(expect.begin(), expect.end(),
                             [](const auto& a, const auto& b) { return a.second > b.second; });
            if (expect.size() > 5) expect.resize(5);
            CHECK(top == expect);
            // the same for a token-set ratio (scored over the whole set, then selected)
            rfa::fuzz::CachedTokenSetRatio<char> tset(choices[3]);
            const auto ttop = tset.extract_topk(set, 3, 30.0);
            std::vector<std::pair<int64_t, double>> texpect;
            const auto tall = tset.similarity_many(set, 30.0);
            for (size_t c = 0; c < choices.size(); ++c)
                if (tall[c] >= 30.0) texpect.emplace_back((int64_t)c, tall[c]);
            std::stable_sort(texpect.begin(), texpect.end(),
                             [](const auto& a, const auto& b) { return a.second > b.second; });
            if (texpect.size() > 3) texpect.resize(3);
            CHECK(ttop == texpect);
            CHECK(!ttop.empty() && ttop[0].second == 100.0);
        }

        // ---- experimental::Multi* ----
        {
            rfa::experimental::MultiLevenshtein<8> multi(3);
            CHECK(multi.result_count() == 32);
            multi.insert(std::string("abc"));
            multi.insert(std::string(""));
            multi.insert(std::string("abcdefgh"));
            std::vector<size_t> scores(multi.result_count());
            multi.distance(scores.data(), scores.size(), std::string("abd"));
            CHECK(scores[0] == 1 && scores[1] == 3 && scores[2] == 5);
            CHECK(scores[31] == 3); // padding lanes score the empty string
            CHECK(throws_invalid([&] { multi.insert(std::string("x")); }));
            CHECK(throws_invalid([&] { multi.distance(scores.data(), 31, std::string("abd")); }));
            CHECK(throws_invalid([] { rfa::experimental::MultiLevenshtein<8> bad(2, rfa::LevenshteinWeightTable{1, 1, 3}); }));
            std::vector<double> ns(multi.result_count());
            multi.normalized_similarity(ns.data(), ns.size(), std::string("abd"));
            rfa::CachedLevenshtein<char> c0("abc");
            CHECK(near(ns[0], 1.0 - 1.0 / 3.0));
            (void)c0;
        }
        {
            rfa::fuzz::experimental::MultiRatio<16> mr(2);
            mr.insert(std::string("new york mets"));
            mr.insert(std::string("atlanta braves"));
            std::vector<double> s(mr.result_count());
            mr.similarity(s.data(), s.size(), std::string("the wonderful new york mets"));
            CHECK(near(s[0], 65.0));
        }
        {
            // tests-Hamming.cpp:99-108
            rfa::CachedHamming<char> h(std::string("aaaa"));
            CHECK(h.distance(std::string("aaaa")) == 0);
            CHECK(h.distance(std::string("abaa")) == 1);
            CHECK(h.distance(std::string("aaba")) == 1);
            CHECK(rfa::CachedHamming<char>(std::string("abaa")).distance(std::string("aaba")) == 2);
            CHECK(h.distance(std::string("aaaaa")) == 1);
            CHECK(rfa::CachedHamming<char>(std::string("aaaaa")).distance(std::string("aaaa")) == 1);
            CHECK(h.similarity(std::string("abaa")) == 3);
            CHECK(near(h.normalized_distance(std::string("abab")), 0.5));
            rfa::CachedHamming<char> strict(std::string("aaaa"), false);
            CHECK(strict.distance(std::string("abaa")) == 1);
            CHECK(throws_invalid([&] { (void)strict.distance(std::string("aaaaa")); }));
            rfa::CachedPrefix<char> pre(std::string("prefix_sample"));
            CHECK(pre.similarity(std::string("prefix_other")) == 7);
            CHECK(pre.similarity(std::string("prefix_other"), 8) == 0);
            CHECK(pre.distance(std::string("prefix_other")) == 6);
            rfa::CachedPostfix<char> post(std::string("sample_postfix"));
            CHECK(post.similarity(std::string("other_postfix")) == 8);
            CHECK(near(post.normalized_similarity(std::string("other_postfix")), 8.0 / 14.0));
            // batch equals one-pair calls
            std::mt19937 rng(5);
            std::vector<std::string> cs;
            for (int i = 0; i < 300; ++i) cs.push_back(random_string(rng, 0, 40));
            rfa::Choices ch(cs);
            const std::string q = random_string(rng, 10, 30);
            rfa::CachedHamming<char> hq(q);
            rfa::CachedPrefix<char> pq(q);
            rfa::CachedPostfix<char> sq(q);
            auto hd = hq.distance_many(ch);
            auto pd = pq.similarity_many(ch);
            auto sd = sq.similarity_many(ch);
            for (size_t i = 0; i < cs.size(); ++i) {
                CHECK(hd[i] == hq.distance(cs[i]));
                CHECK(pd[i] == pq.similarity(cs[i]));
                CHECK(sd[i] == sq.similarity(cs[i]));
            }
        }
        {
            // free functions: known answers (tests-Levenshtein.cpp:144-179, tests-Indel.cpp:204-215,
            // tests-OSA.cpp:62-97, tests-fuzz.cpp:63-99, tests-Hamming.cpp:99-108)
            CHECK(rfa::levenshtein_distance(std::string("South Korea"), std::string("North Korea")) == 2);
            CHECK(rfa::levenshtein_distance(std::string("South Korea"), std::string("North Korea"), {1, 1, 2}) == 4);
            CHECK(rfa::levenshtein_distance(std::string("South Korea"), std::string("North Korea"), {1, 1, 2}, 2) == 3);
            CHECK(rfa::levenshtein_distance("aabc", "cccd", {1, 1, 1}, 2) == 3);
            const std::string sk = "South Korea", nk = "North Korea";
            CHECK(rfa::levenshtein_distance(sk.begin(), sk.end(), nk.begin(), nk.end()) == 2);
            CHECK(rfa::levenshtein_similarity(sk, nk) == 9);
            CHECK(near(rfa::levenshtein_normalized_distance(sk, nk), 2.0 / 11.0));
            CHECK(near(rfa::levenshtein_normalized_similarity(sk, nk), 9.0 / 11.0));
            CHECK(rfa::levenshtein_normalized_similarity(sk, nk, {1, 1, 1}, 0.9) == 0.0);
            CHECK(rfa::indel_distance(std::string("aaaa"), std::string("bbbb")) == 8);
            CHECK(rfa::indel_similarity(std::string("aaaa"), std::string("aaaa")) == 8);
            CHECK(rfa::lcs_seq_similarity(std::string("aaaa"), std::string("aaab")) == 3);
            CHECK(rfa::lcs_seq_distance(std::string("aaaa"), std::string("aaab")) == 1);
            CHECK(rfa::osa_distance(std::string("CA"), std::string("ABC")) == 3);
            CHECK(rfa::osa_distance(std::string("CA"), std::string("AC")) == 1);
            CHECK(rfa::osa_distance(std::string("aaaa"), std::string(""), 1) == 2);
            CHECK(near(rfa::jaro_winkler_similarity(std::string("0"), std::string("00")), 0.85));
            CHECK(near(rfa::jaro_similarity(std::string("00"), std::string("00")), 1.0));
            CHECK(near(rfa::jaro_winkler_distance(std::string("0"), std::string("00"), 0.1), 0.15));
            CHECK(rfa::hamming_distance(std::string("abaa"), std::string("aaba")) == 2);
            CHECK(rfa::hamming_distance(std::string("aaaa"), std::string("aaaaa")) == 1);
            CHECK(throws_invalid([] { (void)rfa::hamming_distance(std::string("aaaa"), std::string("aaaaa"), false); }));
            CHECK(rfa::prefix_similarity(std::string("prefix_sample"), std::string("prefix_other")) == 7);
            CHECK(rfa::postfix_similarity(std::string("sample_postfix"), std::string("other_postfix")) == 8);
            CHECK(rfa::experimental::damerau_levenshtein_distance(std::string("CA"), std::string("ABC")) == 2);
            CHECK(near(rfa::fuzz::ratio(std::string("new york mets"), std::string("the wonderful new york mets")), 65.0));
            CHECK(near(rfa::fuzz::ratio(std::string("new york mets"), std::string("the wonderful new york mets"), 70.0),
                       0.0));
            CHECK(near(rfa::fuzz::partial_ratio(std::string("the wonderful new york mets"), std::string("new york mets")),
                       100.0));
            CHECK(near(rfa::fuzz::token_sort_ratio(std::string("new york mets"), std::string("mets york new")), 100.0));
            const std::string s4 = "new york mets vs atlanta braves", s5 = "atlanta braves vs new york mets";
            CHECK(near(rfa::fuzz::token_set_ratio(s4, s5), 100.0));
            CHECK(near(rfa::fuzz::token_ratio(s4, s5), 100.0));
            CHECK(near(rfa::fuzz::partial_token_set_ratio(s4, std::string("new york city mets - atlanta braves")), 100.0));
            CHECK(near(rfa::fuzz::QRatio(std::string("test"), std::string("")), 0.0));
            CHECK(near(rfa::fuzz::partial_token_sort_ratio(std::string(""), std::string("")), 100.0));
            CHECK(near(rfa::fuzz::partial_token_ratio(std::string(""), std::string("")), 100.0));
            // WRatio (tests-fuzz.cpp:101-135): equal, partial (x 0.9), misordered (x 0.95), issue 452
            const std::string s1 = "new york mets", s3 = "the wonderful new york mets";
            CHECK(near(rfa::fuzz::WRatio(s1, s1), 100.0));
            CHECK(near(rfa::fuzz::WRatio(s1, s3), 90.0));
            CHECK(near(rfa::fuzz::WRatio(s4, s5), 95.0));
            std::string hb = "hello";
            for (int i = 0; i < 7; ++i) hb += "abcde";
            CHECK(near(rfa::fuzz::WRatio(std::string("hello"), hb), 90.0));
            CHECK(near(rfa::fuzz::WRatio(std::string(""), std::string("")), 0.0));
            CHECK(near(rfa::fuzz::CachedWRatio<char>(s1).similarity(s3, 90.5), 0.0));
            {
                rfa::fuzz::CachedWRatio<char> wr(s4);
                CHECK(near(wr.similarity(s5), rfa::fuzz::WRatio(s4, s5)));
            }
            // free functions on C strings and wide strings
            CHECK(rfa::levenshtein_distance("kitten", "sitting") == 3);
            CHECK(rfa::indel_distance(std::u32string(U"k\u00e4se"), std::u32string(U"kase")) == 2);
        }
        {
            // pairs_*: one launch over many pairs == the per-pair free functions
            std::mt19937 rng(11);
            std::vector<std::string> as, bs;
            std::vector<int32_t> mets;
            for (int i = 0; i < 400; ++i) {
                as.push_back(random_string(rng, 0, i % 50 == 0 ? 700 : 90));
                bs.push_back(random_string(rng, 0, 90));
                mets.push_back(i % 4);
            }
            const rfa::Choices A(as), B(bs);
            const auto pd = rfa::pairs_distance(A, B, mets);
            const auto pdc = rfa::pairs_distance(A, B, mets, 7);
            const auto ps = rfa::pairs_similarity(A, B, mets);
            const auto pnd = rfa::pairs_normalized_distance(A, B, mets, 0.4);
            const auto pns = rfa::pairs_normalized_similarity(A, B, mets, 0.6);
            const auto lev = rfa::pairs_distance(A, B, rfa::PairMetric::Levenshtein);
            const auto plain = rfa::pairs_distance(A, B, mets, std::numeric_limits<size_t>::max(), RFA_SCHEDULE_PLAIN);
            for (size_t i = 0; i < as.size(); i += 3) {
                size_t d, dc, sim;
                double nd, ns;
                switch (mets[i]) {
                case 0:
                    d = rfa::levenshtein_distance(as[i], bs[i]);
                    dc = rfa::levenshtein_distance(as[i], bs[i], {1, 1, 1}, 7);
                    sim = rfa::levenshtein_similarity(as[i], bs[i]);
                    nd = rfa::levenshtein_normalized_distance(as[i], bs[i], {1, 1, 1}, 0.4);
                    ns = rfa::levenshtein_normalized_similarity(as[i], bs[i], {1, 1, 1}, 0.6);
                    break;
                case 1:
                    d = rfa::indel_distance(as[i], bs[i]);
                    dc = rfa::indel_distance(as[i], bs[i], 7);
                    sim = rfa::indel_similarity(as[i], bs[i]);
                    nd = rfa::indel_normalized_distance(as[i], bs[i], 0.4);
                    ns = rfa::indel_normalized_similarity(as[i], bs[i], 0.6);
                    break;
                case 2:
                    d = rfa::lcs_seq_distance(as[i], bs[i]);
                    dc = rfa::lcs_seq_distance(as[i], bs[i], 7);
                    sim = rfa::lcs_seq_similarity(as[i], bs[i]);
                    nd = rfa::lcs_seq_normalized_distance(as[i], bs[i], 0.4);
                    ns = rfa::lcs_seq_normalized_similarity(as[i], bs[i], 0.6);
                    break;
                default:
                    d = rfa::osa_distance(as[i], bs[i]);
                    dc = rfa::osa_distance(as[i], bs[i], 7);
                    sim = rfa::osa_similarity(as[i], bs[i]);
                    nd = rfa::osa_normalized_distance(as[i], bs[i], 0.4);
                    ns = rfa::osa_normalized_similarity(as[i], bs[i], 0.6);
                    break;
                }
                CHECK(pd[i] == d);
                CHECK(pdc[i] == dc);
                CHECK(ps[i] == sim);
                CHECK(pnd[i] == nd);
                CHECK(pns[i] == ns);
                CHECK(plain[i] == d);
                CHECK(lev[i] == rfa::levenshtein_distance(as[i], bs[i]));
            }
            CHECK(throws_invalid([&] { (void)rfa::pairs_distance(A, rfa::Choices(std::vector<std::string>{"x"}), mets); }));
        }
        {
            // Server: every answer equals extract_topk of the same query on the same set
            std::mt19937 rng(13);
            std::vector<std::string> cs;
            for (int i = 0; i < 2000; ++i) cs.push_back(random_string(rng, 4, 40));
            const rfa::Choices set(cs);
            rfa::ServerConfig cfg;
            cfg.metric = RFA_RATIO;
            cfg.kind = RFA_SIMILARITY;
            cfg.score_cutoff = 50.0;
            cfg.k = 5;
            rfa::Server srv(set, cfg);
            std::vector<std::string> qs;
            std::vector<uint64_t> tickets;
            for (int i = 0; i < 24; ++i) {
                qs.push_back(random_string(rng, 4, 40));
                tickets.push_back(srv.submit(qs.back()));
            }
            for (size_t i = 0; i < qs.size(); ++i) {
                const auto got = srv.wait(tickets[i]);
                const auto exp = rfa::fuzz::CachedRatio<char>(qs[i]).extract_topk(set, 5, 50.0);
                CHECK(got == exp);
            }
            const auto self = srv.query(cs[17]);
            CHECK(!self.empty() && self[0].second == 100.0);
            CHECK(srv.stats().requests >= 25);
            CHECK(throws_invalid([&] { (void)srv.submit(std::u32string(U"\u4e2d")); }));
        }
        {
            // Multi* with one cutoff per inserted string == one cutoff at a time
            std::mt19937 rng(17);
            rfa::experimental::MultiLevenshtein<32> ml(5);
            rfa::experimental::MultiIndel<32> mi(5);
            std::vector<std::string> qs;
            for (int i = 0; i < 5; ++i) {
                qs.push_back(random_string(rng, 1, 32));
                ml.insert(qs.back());
                mi.insert(qs.back());
            }
            std::vector<std::string> ts;
            for (int i = 0; i < 50; ++i) ts.push_back(random_string(rng, 0, 40));
            const rfa::Choices tc(ts);
            const std::vector<size_t> cuts = {0, 1, 3, 6, 40};
            const auto dm = ml.distance_many(tc, cuts);
            const auto sm = mi.similarity_many(tc, cuts);
            const size_t rc = ml.result_count();
            std::vector<size_t> one(rc);
            for (size_t t = 0; t < ts.size(); ++t)
                for (size_t i = 0; i < qs.size(); ++i) {
                    ml.distance(one.data(), rc, ts[t], cuts[i]);
                    CHECK(dm[t * rc + i] == one[i]);
                    mi.similarity(one.data(), mi.result_count(), ts[t], cuts[i]);
                    CHECK(sm[t * mi.result_count() + i] == one[i]);
                }
            CHECK(throws_invalid([&] { (void)ml.distance_many(tc, std::vector<size_t>{1, 2}); }));
        }
        {
            rfa::experimental::MultiIndel<64> mi(1);
            mi.insert(std::string(64, 'a'));
            std::vector<size_t> s(mi.result_count());
            mi.distance(s.data(), s.size(), std::string(64, 'b'));
            CHECK(s[0] == 128);
        }
    }
    catch (const rfa::Error& e) {
        if (e.status == RFA_ERR_NO_DEVICE) {
            std::printf("no GPU device: %s\n", e.what());
            return 77;
        }
        std::printf("device error: %s\n", e.what());
        return 1;
    }
    catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    std::printf("%s: %d checks, %d failures\n", g_fail ? "FAIL" : "OK", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
