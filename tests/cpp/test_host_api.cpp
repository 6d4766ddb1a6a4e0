// C++ host API test: the reference's scalar known-answer tests
// (test/unit/affine/*_scalar*_test.cpp, test/unit/configuration/*.cpp), written the way the
// reference's tests configure and call the aligner, plus one-to-one bulk and one-to-many calls
// checked against the single-pair call, and FASTA / FASTQ records read for compute().
//
//   test_host_api --config   checks the configuration layer and the record reader (no GPU needed)
//   test_host_api            runs everything on device 0
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <string_view>
#include <vector>

#include <pairwise_aligner/amd/fastx.hpp>
#include <pairwise_aligner/configuration/configure_aligner.hpp>
#include <pairwise_aligner/configuration/gap_model_affine.hpp>
#include <pairwise_aligner/configuration/initial.hpp>
#include <pairwise_aligner/configuration/method_global.hpp>
#include <pairwise_aligner/configuration/method_local.hpp>
#include <pairwise_aligner/configuration/score_model_matrix.hpp>
#include <pairwise_aligner/configuration/score_model_unitary.hpp>
#include <pairwise_aligner/configuration/score_model_unitary_simd.hpp>
#include <pairwise_aligner/score_model/substitution_matrix.hpp>

using namespace std::literals;
namespace aligner = seqan::pairwise_aligner;

static int failures = 0;
#define EXPECT_EQ(a, b)                                                                                   \
    do {                                                                                                  \
        auto va_ = (a);                                                                                   \
        auto vb_ = (b);                                                                                   \
        if (!(va_ == vb_)) {                                                                              \
            std::printf("FAIL %s:%d: %s == %lld, expected %lld\n", __FILE__, __LINE__, #a, (long long)va_, \
                        (long long)vb_);                                                                  \
            ++failures;                                                                                   \
        }                                                                                                 \
    } while (0)

inline constexpr auto standard = aligner::cfg::method_global(aligner::cfg::gap_model_affine(-10, -1),
                                                             aligner::cfg::leading_end_gap{},
                                                             aligner::cfg::trailing_end_gap{});
inline constexpr auto overlap = aligner::cfg::method_global(
    aligner::cfg::gap_model_affine(-10, -1),
    aligner::cfg::leading_end_gap{aligner::cfg::end_gap::free, aligner::cfg::end_gap::free},
    aligner::cfg::trailing_end_gap{aligner::cfg::end_gap::free, aligner::cfg::end_gap::free});
inline constexpr auto semi_first = aligner::cfg::method_global(
    aligner::cfg::gap_model_affine(-10, -1), aligner::cfg::leading_end_gap{.first_column = aligner::cfg::end_gap::free},
    aligner::cfg::trailing_end_gap{.last_column = aligner::cfg::end_gap::free});
inline constexpr auto semi_second = aligner::cfg::method_global(
    aligner::cfg::gap_model_affine(-10, -1), aligner::cfg::leading_end_gap{.first_row = aligner::cfg::end_gap::free},
    aligner::cfg::trailing_end_gap{.last_row = aligner::cfg::end_gap::free});
inline constexpr auto local = aligner::cfg::method_local(aligner::cfg::gap_model_affine(-10, -1));

template <typename config_t>
static int32_t score_of(config_t const &config, std::string_view s1, std::string_view s2)
{
    auto a = aligner::cfg::configure_aligner(config);
    return a.compute(s1, s2).score();
}

static void config_checks()
{
    auto a = aligner::cfg::configure_aligner(aligner::cfg::method_global(
        aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary(4, -5), -10, -1),
        aligner::cfg::leading_end_gap{.first_column = aligner::cfg::end_gap::penalised,
                                      .first_row = aligner::cfg::end_gap::free},
        aligner::cfg::trailing_end_gap{.last_column = aligner::cfg::end_gap::penalised,
                                       .last_row = aligner::cfg::end_gap::free}));
    const pa_config &c = a.config();
    EXPECT_EQ(c.method, PA_METHOD_GLOBAL);
    EXPECT_EQ(c.first_column_free, 0);
    EXPECT_EQ(c.first_row_free, 1);
    EXPECT_EQ(c.last_column_free, 0);
    EXPECT_EQ(c.last_row_free, 1);
    EXPECT_EQ(c.gap_open, -10);
    EXPECT_EQ(c.gap_extension, -1);
    EXPECT_EQ(c.match, 4);
    EXPECT_EQ(c.mismatch, -5);
    EXPECT_EQ(c.lane_policy, PA_LANES_AUTO);
    auto m = aligner::cfg::configure_aligner(aligner::cfg::score_model_matrix(local, aligner::blosum62_extended<>));
    EXPECT_EQ(m.config().method, PA_METHOD_LOCAL);
    EXPECT_EQ(m.config().matrix_dim, 27);
    EXPECT_EQ(int(m.symbols()[26]), int('*'));
    EXPECT_EQ(m.matrix()[26 * 27 + 26], 1);
    auto s = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary_simd(standard, (int16_t)4, (int16_t)-5));
    EXPECT_EQ(s.config().lane_policy, PA_LANES_INT16);
    auto d = aligner::cfg::configure_aligner(aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary(1, -1), -2, -1));
    EXPECT_EQ(d.config().method, PA_METHOD_GLOBAL);   // default method

    // chains that start from cfg::initial or from a method without predecessor
    // (initial.hpp:74, method_global.hpp:137-140, method_local.hpp:129-132)
    using aligner::cfg::end_gap;
    auto l = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(
        aligner::cfg::gap_model_affine(aligner::cfg::method_local(), -10, -1), 4, -5));
    EXPECT_EQ(l.config().method, PA_METHOD_LOCAL);
    EXPECT_EQ(l.config().gap_open, -10);
    EXPECT_EQ(l.config().match, 4);
    auto g = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(
        aligner::cfg::gap_model_affine(aligner::cfg::method_global(aligner::cfg::leading_end_gap{end_gap::free, end_gap::penalised},
                                                                   aligner::cfg::trailing_end_gap{end_gap::penalised, end_gap::free}),
                                       -10, -1),
        4, -5));
    EXPECT_EQ(g.config().method, PA_METHOD_GLOBAL);
    EXPECT_EQ(g.config().first_column_free, 1);
    EXPECT_EQ(g.config().first_row_free, 0);
    EXPECT_EQ(g.config().last_column_free, 0);
    EXPECT_EQ(g.config().last_row_free, 1);
    auto i = aligner::cfg::configure_aligner(
        aligner::cfg::method_local(aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary(aligner::cfg::initial, 2, -1), -1, -1)));
    EXPECT_EQ(i.config().method, PA_METHOD_LOCAL);
    EXPECT_EQ(i.config().mismatch, -1);
    static_assert(std::is_same_v<aligner::cfg::initial_t, std::remove_cvref_t<decltype(aligner::cfg::initial)>>);
}

static void known_answers()
{
    using aligner::cfg::score_model_unitary;
    using aligner::cfg::score_model_matrix;
    // global_standard_affine_scalar_test.cpp:34-53
    EXPECT_EQ(score_of(score_model_unitary(standard, 4, -5), "ACGTGACTGACACTACGACT"sv, "ACGTGACTGACACTACGACT"sv), 80);
    EXPECT_EQ(score_of(score_model_unitary(standard, 4, -5), "AAAAAAAAAA"sv, "TTTTTTTTTT"sv), -40);
    EXPECT_EQ(score_of(score_model_unitary(standard, 4, -5), "AGATCGACTAGCGAGCTACGAGCTAGC"sv, "AGACGATCGACGAGCGACTACGTACGA"sv), 26);
    // global_overlap_affine_scalar_test.cpp:35-54
    EXPECT_EQ(score_of(score_model_unitary(overlap, 4, -5), "ACGTGACTGACACTACGACT"sv, "ACGTGACTGACACTACGACT"sv), 80);
    EXPECT_EQ(score_of(score_model_unitary(overlap, 4, -5), "TTTTTCTGACACACT"sv, "CTGCACACGACGGGGGG"sv), 16);
    EXPECT_EQ(score_of(score_model_unitary(overlap, 4, -5), "CTGCACACGACGGGGGG"sv, "TTTTTCTGACACACT"sv), 16);
    // global_semi_first_affine_scalar_test.cpp:35-54
    EXPECT_EQ(score_of(score_model_unitary(semi_first, 4, -5), "TTTTTACGTATGTCCCCC"sv, "ACGTAAAACGT"sv), 10);
    EXPECT_EQ(score_of(score_model_unitary(semi_first, 4, -5), "AAGACTACGGGGGG"sv, "ACGAGGCTAC"sv), 11);
    EXPECT_EQ(score_of(score_model_unitary(semi_first, 4, -5), "GGGGGGAAGACTAC"sv, "ACGAGGCTAC"sv), 11);
    // global_semi_second_affine_scalar_test.cpp:35-54
    EXPECT_EQ(score_of(score_model_unitary(semi_second, 4, -5), "ACGTAAAACGT"sv, "TTTTTACGTATGTCCCCC"sv), 10);
    EXPECT_EQ(score_of(score_model_unitary(semi_second, 4, -5), "ACGAGGCTAC"sv, "AAGACTACGGGGGG"sv), 11);
    EXPECT_EQ(score_of(score_model_unitary(semi_second, 4, -5), "ACGAGGCTAC"sv, "GGGGGGAAGACTAC"sv), 11);
    // global_standard_affine_scalar_matrix_test.cpp:32-58
    EXPECT_EQ(score_of(score_model_matrix(standard, aligner::blosum62_standard<>), "ACKLMNPQRRTVWYNMPQHIK"sv, "ACKLMNPQRRTVWYNMPQHIK"sv), 122);
    EXPECT_EQ(score_of(score_model_matrix(standard, aligner::blosum62_standard<>), "AAAAAAAAAAAAAAAAAAAA"sv, "CDEFGHIKLMNPQRSTVWLY"sv), -21);
    EXPECT_EQ(score_of(score_model_matrix(standard, aligner::blosum62_standard<>), "FNQSAEYPDISHCGVMQLKWRATLGT"sv, "EIKSDVLLHRWSMKNPGNILMIDVGMQVAESYFAT"sv), -26);
    EXPECT_EQ(score_of(score_model_matrix(standard, aligner::blosum62_standard<>), "RKFCYMD"sv, "GAYQW"sv), -11);
    // local_affine_scalar_test.cpp:39-105
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "TTTTTTTTTTTTTTT"sv, "TTTTTTTTTTTTTTT"sv), 60);
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "CTGCACACGACGGGGGG"sv, "TTTTTCTGACACACT"sv), 21);
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "AAAAAAAAAA"sv, "TTTTTTTTTT"sv), 0);
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "AAAAATAAAA"sv, "CCCCCTCCCC"sv), 4);
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "AACCGGTTTAACCGGTT"sv, "ACGTCTACGTA"sv), 11);
    EXPECT_EQ(score_of(score_model_unitary(local, 4, -5), "AAAAAATCCCCCC"sv, "CCCCCCTAAAAAA"sv), 24);
    EXPECT_EQ(score_of(score_model_unitary(aligner::cfg::method_local(aligner::cfg::gap_model_affine(-1, -1)), 2, -1),
                       "ATAAGCGTCTCG"sv, "TCATAGAGTTGC"sv), 9);
    EXPECT_EQ(score_of(score_model_matrix(local, aligner::blosum62_extended<>), "ALIGATOR"sv, "GALORA"sv), 13);
    EXPECT_EQ(score_of(score_model_matrix(local, aligner::blosum62_extended<>), "ALIGATOR"sv, ""sv),
              std::numeric_limits<int32_t>::lowest());
    // the same through chains that start from the method rule or from cfg::initial
    EXPECT_EQ(score_of(aligner::cfg::score_model_unitary(aligner::cfg::gap_model_affine(aligner::cfg::method_local(), -10, -1), 4, -5),
                       "TTTTTTTTTTTTTTT"sv, "TTTTTTTTTTTTTTT"sv), 60);
    EXPECT_EQ(score_of(aligner::cfg::score_model_unitary(
                           aligner::cfg::gap_model_affine(
                               aligner::cfg::method_global(
                                   aligner::cfg::leading_end_gap{aligner::cfg::end_gap::free, aligner::cfg::end_gap::free},
                                   aligner::cfg::trailing_end_gap{aligner::cfg::end_gap::free, aligner::cfg::end_gap::free}),
                               -10, -1),
                           4, -5),
                       "TTTTTCTGACACACT"sv, "CTGCACACGACGGGGGG"sv), 16);
    EXPECT_EQ(score_of(aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary(aligner::cfg::initial, 4, -5), -10, -1),
                       "ACGTGACTGACACTACGACT"sv, "ACGTGACTGACACTACGACT"sv), 80);
    // configuration_test.cpp:18-37
    auto proto = aligner::cfg::configure_aligner(aligner::cfg::method_global(
        aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary(4, -5), -10, -1),
        aligner::cfg::leading_end_gap{.first_column = aligner::cfg::end_gap::penalised,
                                      .first_row = aligner::cfg::end_gap::free},
        aligner::cfg::trailing_end_gap{.last_column = aligner::cfg::end_gap::penalised,
                                       .last_row = aligner::cfg::end_gap::free}));
    std::string_view seq{"ACGTGACTGACACTACGACT"};
    EXPECT_EQ(proto.compute(seq, seq).score(), 80);
    // configure_aligner_saturated_test.cpp:18-42 (bulk of one)
    std::vector collection1{seq};
    std::vector collection2{seq};
    auto sat = aligner::cfg::configure_aligner(aligner::cfg::method_global(
        aligner::cfg::gap_model_affine(aligner::cfg::score_model_unitary_simd((int16_t)4, (int16_t)-5), -10, -1),
        aligner::cfg::leading_end_gap{}, aligner::cfg::trailing_end_gap{}));
    EXPECT_EQ(sat.compute(collection1, collection2)[0].score(), 80);
}

static void bulk_calls()
{
    // one-to-one bulk larger than any SIMD width, and one-to-many, against single calls
    std::vector<std::string> s1, s2;
    unsigned x = 12345;
    auto rnd = [&] { x = x * 1103515245u + 12345u; return (x >> 16) & 0x7fff; };
    for (int i = 0; i < 300; ++i) {
        std::string a, b;
        for (int k = 0, n = 1 + rnd() % 120; k < n; ++k) a.push_back("ACGT"[rnd() % 4]);
        for (int k = 0, n = 1 + rnd() % 120; k < n; ++k) b.push_back("ACGT"[rnd() % 4]);
        s1.push_back(a);
        s2.push_back(b);
    }
    auto a = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(local, 4, -5));
    auto bulk = a.compute(s1, s2);
    EXPECT_EQ(bulk.size(), s1.size());
    for (std::size_t i = 0; i < s1.size(); i += 37) {
        EXPECT_EQ(bulk[i].score(), a.compute(s1[i], s2[i]).score());
        EXPECT_EQ(bulk[i].sequence1() == s1[i], true);
    }
    auto many = a.compute(s1[0], s2);
    for (std::size_t i = 0; i < s2.size(); i += 41) {
        EXPECT_EQ(many[i].score(), a.compute(s1[0], s2[i]).score());
        EXPECT_EQ(many[i].sequence1() == s1[0], true);
    }
}

// the two halves of a vector of cells
static std::vector<int32_t> scores_of(std::vector<aligner::dp_cell> const &v)
{
    std::vector<int32_t> out;
    for (auto const &c : v) out.push_back(c.score());
    return out;
}
static std::vector<int32_t> gaps_of(std::vector<aligner::dp_cell> const &v)
{
    std::vector<int32_t> out;
    for (auto const &[score, gap] : v) out.push_back(gap);
    return out;
}

// compute(seq1, seq2, first_dp_column, first_dp_row): the final DP column and row.  Expected
// values worked by hand from the recurrence (+4/-5, gap (-10,-1), global): H(4, j) of ACGT x AGT is
// -14 -9 -13 1 and H(i, 3) is -13 -8 -12 -6 1.  The gap halves: below row 4, V(5, j) =
// max over i of H(i, j) - 10 - (5 - i), e.g. V(5, 1) = H(1, 1) - 14 = -10; right of column 3,
// W(i, 4) = max over j of H(i, j) - 10 - (4 - j), e.g. W(4, 4) = H(4, 3) - 11 = -10; entry 0 of
// each vector -11.
void dp_vectors()
{
    auto a = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(standard, 4, -5));
    std::vector<int> unused_column, unused_row;
    auto r = a.compute(std::string{"ACGT"}, std::string{"AGT"}, unused_column, unused_row);
    EXPECT_EQ(r.score(), 1);
    EXPECT_EQ(scores_of(r.dp_row()) == (std::vector<int32_t>{-14, -9, -13, 1}), true);
    EXPECT_EQ(scores_of(r.dp_column()) == (std::vector<int32_t>{-13, -8, -12, -6, 1}), true);
    EXPECT_EQ(gaps_of(r.dp_row()) == (std::vector<int32_t>{-11, -10, -14, -10}), true);
    EXPECT_EQ(gaps_of(r.dp_column()) == (std::vector<int32_t>{-11, -9, -13, -15, -10}), true);
    EXPECT_EQ(std::get<1>(r.dp_row()[1]), -10);
    EXPECT_EQ(r.dp_column()[4].second, -10);
    auto e = a.compute(std::string{""}, std::string{"CC"}, unused_column, unused_row);
    EXPECT_EQ(scores_of(e.dp_row()) == (std::vector<int32_t>{0, -11, -12}), true);
    EXPECT_EQ(scores_of(e.dp_column()) == (std::vector<int32_t>{-12}), true);
    EXPECT_EQ(gaps_of(e.dp_row()) == (std::vector<int32_t>{-11, -22, -23}), true);
    EXPECT_EQ(gaps_of(e.dp_column()) == (std::vector<int32_t>{-11}), true);

    // bulk and one-to-many forms (interface_one_to_one_bulk.hpp:73-100,
    // interface_one_to_many_bulk.hpp:67-77; aligner_result_bulk.hpp:43-51): each result carries its
    // own pair's vectors, equal to the single-pair call's
    std::vector<std::string> s1{"ACGT", "", "ACGTACGGT", "TTAG", "GATTACA"}, s2{"AGT", "CC", "AGGT", "", "GATCA"};
    for (auto const &cfg_a : {a, aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(local, 4, -5))}) {
        auto b = cfg_a;
        auto bulk = b.compute(s1, s2, unused_column, unused_row);
        EXPECT_EQ(bulk.size(), s1.size());
        for (std::size_t k = 0; k < s1.size(); ++k) {
            auto one = b.compute(s1[k], s2[k], unused_column, unused_row);
            EXPECT_EQ(bulk[k].score(), one.score());
            EXPECT_EQ(bulk[k].dp_column() == one.dp_column(), true);
            EXPECT_EQ(bulk[k].dp_row() == one.dp_row(), true);
            EXPECT_EQ(bulk[k].dp_column().size(), s1[k].size() + 1);
        }
        auto many = b.compute(s1[0], s2, unused_column, unused_row);
        EXPECT_EQ(many.size(), s2.size());
        for (std::size_t k = 0; k < s2.size(); ++k) {
            auto one = b.compute(s1[0], s2[k], unused_column, unused_row);
            EXPECT_EQ(many[k].score(), one.score());
            EXPECT_EQ(many[k].dp_column() == one.dp_column(), true);
            EXPECT_EQ(many[k].dp_row() == one.dp_row(), true);
        }
    }
    auto bulk = a.compute(s1, s2, unused_column, unused_row);
    EXPECT_EQ(scores_of(bulk[0].dp_row()) == (std::vector<int32_t>{-14, -9, -13, 1}), true);
    EXPECT_EQ(scores_of(bulk[0].dp_column()) == (std::vector<int32_t>{-13, -8, -12, -6, 1}), true);
    EXPECT_EQ(gaps_of(bulk[0].dp_row()) == (std::vector<int32_t>{-11, -10, -14, -10}), true);
    auto plain = a.compute(s1, s2);
    EXPECT_EQ(plain[0].dp_column().empty() && plain[0].dp_row().empty(), true);
}

// The reference's FASTQ fixture (test/data/in.fastq), as text and as tests/golden/in.fastq.
static constexpr std::string_view in_fastq =
    "@seq1\nACGTTTGATTCGCG\n+\n!**55CCF>>>>>C\n@seq2\nTCGGGGGATTCGCG\n+\n!**65CCF>CF>>C\n";

static void fastx_checks(std::string const &golden_dir)
{
    auto r = aligner::fastx_records::parse(in_fastq);
    EXPECT_EQ(r.size(), 2u);
    EXPECT_EQ(r.is_fastq(), true);
    EXPECT_EQ(r[0] == "ACGTTTGATTCGCG"sv && r[1] == "TCGGGGGATTCGCG"sv, true);
    EXPECT_EQ(r.name(0) == "seq1"sv && r.name(1) == "seq2"sv, true);
    EXPECT_EQ(r.quality(1) == "!**65CCF>CF>>C"sv, true);
    auto f = aligner::fastx_records::read(golden_dir + "in.fastq");
    EXPECT_EQ(f.size() == 2 && f[0] == r[0] && f[1] == r[1], true);
    auto fa = aligner::fastx_records::parse(">a desc\nacg\ntt\n\n>b\n", true);
    EXPECT_EQ(fa.size() == 2 && fa[0] == "ACGTT"sv && fa[1].empty() && !fa.is_fastq(), true);
    bool threw = false;
    try {
        aligner::fastx_records::parse("@r\nACGT\n+\nII\n");
    } catch (std::runtime_error const &) {
        threw = true;
    }
    EXPECT_EQ(threw, true);
}

// Records read with fastx_records go straight into compute().
static void fastx_compute()
{
    auto r = aligner::fastx_records::parse(in_fastq);
    auto a = aligner::cfg::configure_aligner(aligner::cfg::score_model_unitary(standard, 4, -5));
    auto bulk = a.compute(std::vector<std::string_view>{r[0]}, std::vector<std::string_view>{r[1]});
    EXPECT_EQ(bulk[0].score(), a.compute(std::string{r[0]}, std::string{r[1]}).score());
}

int main(int argc, char **argv)
{
    std::string golden_dir = argv[0];
    golden_dir = golden_dir.substr(0, golden_dir.rfind('/') + 1) + "../golden/";
    config_checks();
    fastx_checks(golden_dir);
    if (!(argc > 1 && std::strcmp(argv[1], "--config") == 0)) {
        known_answers();
        bulk_calls();
        dp_vectors();
        fastx_compute();
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
