// host_api.hpp -- header-only C++20 host layer of the MI355X aligner.
//
// Keeps the configuration and alignment-call API of rrahn/pairwise_aligner (namespace
// seqan::pairwise_aligner, inline namespace v1) so that code written against the reference
// compiles against this tree:
//
//   auto aligner = cfg::configure_aligner(
//       cfg::method_global(cfg::gap_model_affine(cfg::score_model_unitary(4, -5), -10, -1),
//                          cfg::leading_end_gap{}, cfg::trailing_end_gap{}));
//   aligner.compute(seq1, seq2).score();                  // one pair
//   for (auto r : aligner.compute(bulk1, bulk2)) r.score();   // one-to-one bulk, any size
//   for (auto r : aligner.compute(seq1, bulk2)) r.score();    // one-to-many
//
// Rules are literal types, so `inline constexpr auto base_config = cfg::method_global(...)`
// works as in the reference's tests.  configure_aligner walks the chain like
// configure_aligner.hpp:165-185 (exactly one score model and one gap model, checked at
// compile time; the method is optional and defaults to global with penalised end gaps).
// compute() crosses into the device library through the C ABI of pa_device.h; it throws
// std::runtime_error with the library's message on failure (the reference has no error path
// for these cases; its assertions are debug-only).
//
// Differences that callers can observe, all supersets of the reference's behaviour:
//   * bulk size is unbounded (the reference asserts <= the SIMD lane count,
//     interface_one_to_one_bulk.hpp:55-57);
//   * score() is int32_t and exact for every score model, including the *_simd ones whose
//     reference lanes wrap on overflow (simd_score_type.hpp:164-192); saturated int16 lanes
//     are recomputed in int32;
//   * dp_column()/dp_row() hold the cells (score, gap state) of the final DP column and row only
//     after a four-argument compute (one pair, one-to-one bulk or one-to-many); they are empty
//     otherwise; a bulk result's dp_column()/dp_row() is its own pair's, not the whole bulk's.
#pragma once

#include <array>
#include <concepts>
#include <cstdint>
#include <memory>
#include <optional>
#include <mutex>
#include <ranges>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <pa_device.h>
#include <pairwise_aligner/score_model/substitution_matrix.hpp>

namespace seqan::pairwise_aligner {
inline namespace v1 {
namespace cfg {

// end_gap_policy.hpp:20-36
enum struct end_gap : bool { penalised, free };

struct leading_end_gap {
    end_gap first_column{end_gap::penalised};
    end_gap first_row{end_gap::penalised};
};

struct trailing_end_gap {
    end_gap last_column{end_gap::penalised};
    end_gap last_row{end_gap::penalised};
};

namespace detail {

enum struct rule_category { score_model, gap_model, method, lanes };

struct initial_rule {};

template <rule_category category_v, typename payload_t, typename predecessor_t>
struct rule {
    static constexpr rule_category category = category_v;
    payload_t payload;
    predecessor_t predecessor;
};

template <typename t> inline constexpr bool is_rule_v = false;
template <> inline constexpr bool is_rule_v<initial_rule> = true;
template <rule_category c, typename p, typename q> inline constexpr bool is_rule_v<rule<c, p, q>> = true;

template <typename t>
concept rule_type = is_rule_v<std::remove_cvref_t<t>>;

struct unitary_payload {
    int32_t match;
    int32_t mismatch;
    int32_t lanes;
};

template <typename matrix_t>
struct matrix_payload {
    matrix_t matrix;
    int32_t lanes;
};

struct gap_payload {
    int32_t open;
    int32_t extension;
};

struct method_payload {
    bool local;
    leading_end_gap leading;
    trailing_end_gap trailing;
};

struct lanes_payload {
    int32_t policy;
};

template <typename rule_t>
constexpr int count_rules(rule_category c)
{
    if constexpr (std::is_same_v<rule_t, initial_rule>)
        return 0;
    else
        return (rule_t::category == c ? 1 : 0) + count_rules<decltype(rule_t::predecessor)>(c);
}

template <typename t>
concept substitution_table = requires(t const &m) {
    { m.size() } -> std::convertible_to<std::size_t>;
    { m[0].first } -> std::convertible_to<char>;
    { m[0].second.size() } -> std::convertible_to<std::size_t>;
};

} // namespace detail

// configuration/initial.hpp:74: the rule every chain starts from.  Every rule below can take it
// (or no predecessor at all) in place of a predecessor rule.
using initial_t = detail::initial_rule;
inline constexpr initial_t initial{};

// score_model_unitary.hpp:209, score_model_unitary_simd.hpp:230,
// score_model_unitary_simd_saturated.hpp:284
struct score_model_unitary_fn {
    int32_t lanes;

    template <detail::rule_type pred_t, std::integral score_t>
    constexpr auto operator()(pred_t pred, score_t match, score_t mismatch) const
    {
        using payload_t = detail::unitary_payload;
        return detail::rule<detail::rule_category::score_model, payload_t, std::remove_cvref_t<pred_t>>{
            payload_t{static_cast<int32_t>(match), static_cast<int32_t>(mismatch), lanes}, pred};
    }

    template <std::integral score_t>
    constexpr auto operator()(score_t match, score_t mismatch) const
    {
        return (*this)(detail::initial_rule{}, match, mismatch);
    }
};

inline constexpr score_model_unitary_fn score_model_unitary{PA_LANES_AUTO};
inline constexpr score_model_unitary_fn score_model_unitary_simd{PA_LANES_INT16};
inline constexpr score_model_unitary_fn score_model_unitary_simd_saturated{PA_LANES_INT16};

// score_model_matrix.hpp:241, score_model_matrix_simd_NxN.hpp:275, score_model_matrix_simd_1xN.hpp:290
struct score_model_matrix_fn {
    int32_t lanes;

    template <detail::rule_type pred_t, detail::substitution_table matrix_t>
    constexpr auto operator()(pred_t pred, matrix_t const &matrix) const
    {
        using payload_t = detail::matrix_payload<matrix_t>;
        return detail::rule<detail::rule_category::score_model, payload_t, std::remove_cvref_t<pred_t>>{
            payload_t{matrix, lanes}, pred};
    }

    template <detail::substitution_table matrix_t>
    constexpr auto operator()(matrix_t const &matrix) const
    {
        return (*this)(detail::initial_rule{}, matrix);
    }
};

inline constexpr score_model_matrix_fn score_model_matrix{PA_LANES_AUTO};
inline constexpr score_model_matrix_fn score_model_matrix_simd_NxN{PA_LANES_INT16};
inline constexpr score_model_matrix_fn score_model_matrix_simd_1xN{PA_LANES_INT16};
inline constexpr score_model_matrix_fn score_model_matrix_simd_saturated_NxN{PA_LANES_INT16};
inline constexpr score_model_matrix_fn score_model_matrix_simd_saturated_1xN{PA_LANES_INT16};

// gap_model_affine.hpp:155: a gap of length k scores open + k * extension
struct gap_model_affine_fn {
    template <detail::rule_type pred_t, std::integral score_t>
    constexpr auto operator()(pred_t pred, score_t open, score_t extension) const
    {
        using payload_t = detail::gap_payload;
        return detail::rule<detail::rule_category::gap_model, payload_t, std::remove_cvref_t<pred_t>>{
            payload_t{static_cast<int32_t>(open), static_cast<int32_t>(extension)}, pred};
    }

    template <std::integral score_t>
    constexpr auto operator()(score_t open, score_t extension) const
    {
        return (*this)(detail::initial_rule{}, open, extension);
    }
};

inline constexpr gap_model_affine_fn gap_model_affine{};

// method_global.hpp:145
struct method_global_fn {
    template <detail::rule_type pred_t>
    constexpr auto operator()(pred_t pred, leading_end_gap leading = {}, trailing_end_gap trailing = {}) const
    {
        using payload_t = detail::method_payload;
        return detail::rule<detail::rule_category::method, payload_t, std::remove_cvref_t<pred_t>>{
            payload_t{false, leading, trailing}, pred};
    }

    // method_global.hpp:137-140: the method as the first rule of a chain
    constexpr auto operator()(leading_end_gap leading, trailing_end_gap trailing) const
    {
        return (*this)(initial, leading, trailing);
    }
};

inline constexpr method_global_fn method_global{};

// method_local.hpp:137, with every end gap free (method_local.hpp:41-49)
struct method_local_fn {
    template <detail::rule_type pred_t>
    constexpr auto operator()(pred_t pred) const
    {
        using payload_t = detail::method_payload;
        return detail::rule<detail::rule_category::method, payload_t, std::remove_cvref_t<pred_t>>{
            payload_t{true, leading_end_gap{end_gap::free, end_gap::free},
                      trailing_end_gap{end_gap::free, end_gap::free}},
            pred};
    }

    // method_local.hpp:129-132: the method as the first rule of a chain
    constexpr auto operator()() const { return (*this)(initial); }
};

inline constexpr method_local_fn method_local{};

// Extension (no reference counterpart): choose the device lane policy (PA_LANES_*).
struct lane_policy_fn {
    template <detail::rule_type pred_t>
    constexpr auto operator()(pred_t pred, int32_t policy) const
    {
        using payload_t = detail::lanes_payload;
        return detail::rule<detail::rule_category::lanes, payload_t, std::remove_cvref_t<pred_t>>{payload_t{policy},
                                                                                                  pred};
    }
};

inline constexpr lane_policy_fn lane_policy{};

} // namespace cfg

namespace detail {

// One device context per process and device, created on first use.
inline std::shared_ptr<pa_ctx> device_context(int device)
{
    static std::mutex mutex;
    static std::vector<std::weak_ptr<pa_ctx>> contexts(64);
    std::lock_guard<std::mutex> lock{mutex};
    if (device < 0 || device >= int(contexts.size())) throw std::runtime_error("invalid device index");
    if (auto ctx = contexts[device].lock()) return ctx;
    pa_ctx *raw = nullptr;
    if (pa_ctx_create(&raw, device) != PA_OK)
        throw std::runtime_error("pairwise_aligner: cannot create a device context (no gfx950 GPU visible?)");
    std::shared_ptr<pa_ctx> ctx{raw, [](pa_ctx *c) { pa_ctx_destroy(c); }};
    contexts[device] = ctx;
    return ctx;
}

template <typename t>
concept sequence = std::ranges::forward_range<t> && std::convertible_to<std::ranges::range_value_t<t>, char> &&
                   !std::ranges::range<std::ranges::range_reference_t<t>>;

template <typename t>
concept sequence_bulk = std::ranges::forward_range<t> && sequence<std::ranges::range_reference_t<t>>;

template <typename seq_t>
void append(std::vector<uint8_t> &buf, std::vector<uint64_t> &off, seq_t const &s)
{
    for (auto &&c : s) buf.push_back(static_cast<uint8_t>(static_cast<char>(c)));
    off.push_back(buf.size());
}

} // namespace detail

// One entry of the final DP column or row: the score H and the gap state, as the reference's
// affine cell (affine/affine_cell.hpp:25-73: a pair with score() = first, read with get<0>/get<1>
// or a structured binding).  The column's gap state is the horizontal one right of the last
// column, the row's the vertical one below the last row; entry 0 of either holds go + ge.
struct dp_cell : std::pair<int32_t, int32_t> {
    using base_t = std::pair<int32_t, int32_t>;
    using score_type = int32_t;
    using base_t::base_t;

    constexpr int32_t &score() noexcept { return this->first; }
    constexpr int32_t const &score() const noexcept { return this->first; }
};

// aligner_result.hpp:33-105: the sequences (as passed: references stay references) and the score.
template <typename sequence1_t, typename sequence2_t>
struct aligner_result {
    sequence1_t _sequence1;
    sequence2_t _sequence2;
    int32_t _score;
    uint8_t _flags;
    // cells of the final DP column (H(i, n), i = 0..m) and row (H(m, j), j = 0..n); filled by
    // compute(seq1, seq2, first_dp_column, first_dp_row), empty otherwise
    std::vector<dp_cell> _dp_column{};
    std::vector<dp_cell> _dp_row{};

    sequence1_t const &sequence1() const noexcept { return _sequence1; }
    sequence2_t const &sequence2() const noexcept { return _sequence2; }
    int32_t const &score() const noexcept { return _score; }
    uint8_t flags() const noexcept { return _flags; }   // PA_PAIR_* bits
    std::vector<dp_cell> const &dp_column() const noexcept { return _dp_column; }
    std::vector<dp_cell> const &dp_row() const noexcept { return _dp_row; }
};

// aligner_result_bulk.hpp:24-70: a view of one pair of a bulk result shared by all its views.
template <typename bulk1_t, typename bulk2_t>
struct bulk_result {
    bulk1_t _sequences1;
    bulk2_t _sequences2;
    std::vector<int32_t> _scores;
    std::vector<uint8_t> _flags;
    // per pair: the final DP column and row (four-argument compute only)
    std::vector<std::vector<dp_cell>> _dp_columns{};
    std::vector<std::vector<dp_cell>> _dp_rows{};
};

template <typename bulk_result_t, bool one_to_many>
class aligner_result_bulk {
    std::shared_ptr<bulk_result_t> _result{};
    std::size_t _index{};

public:
    aligner_result_bulk() = default;
    aligner_result_bulk(std::shared_ptr<bulk_result_t> result, std::size_t index) noexcept :
        _result{std::move(result)}, _index{index} {}

    decltype(auto) sequence1() const noexcept
    {
        if constexpr (one_to_many)
            return (_result->_sequences1);
        else
            return *std::ranges::next(std::ranges::begin(_result->_sequences1), _index);
    }
    decltype(auto) sequence2() const noexcept
    {
        return *std::ranges::next(std::ranges::begin(_result->_sequences2), _index);
    }
    int32_t score() const noexcept { return _result->_scores[_index]; }
    uint8_t flags() const noexcept { return _result->_flags[_index]; }

    // aligner_result_bulk.hpp:43-51.  The reference returns the bulk's whole DP column / row
    // (one SIMD vector per position); here each pair's own final column, cells (H(i, n), gap
    // state), i = 0..m, and row, j = 0..n, filled by the four-argument compute and empty otherwise.
    std::vector<dp_cell> const &dp_column() const & noexcept
    {
        return _result->_dp_columns.empty() ? none() : _result->_dp_columns[_index];
    }
    std::vector<dp_cell> const &dp_row() const & noexcept
    {
        return _result->_dp_rows.empty() ? none() : _result->_dp_rows[_index];
    }

private:
    static std::vector<dp_cell> const &none() noexcept
    {
        static const std::vector<dp_cell> empty{};
        return empty;
    }
};

// The configured aligner: a pa_config plus a device context (interface_one_to_one_*.hpp).
class aligner {
    pa_config _cfg{};
    std::vector<int32_t> _matrix{};
    std::vector<uint8_t> _symbols{};
    int _device{0};
    std::shared_ptr<pa_ctx> _ctx{};
    std::optional<pa_tuning> _tuning{};

    pa_config device_config()
    {
        if (!_ctx) _ctx = detail::device_context(_device);
        pa_config cfg = _cfg;
        if (!_matrix.empty()) {
            cfg.matrix = _matrix.data();
            cfg.matrix_symbols = _symbols.data();
        }
        if (_tuning) cfg.tuning = &*_tuning;
        return cfg;
    }

    void check(pa_status st)
    {
        if (st != PA_OK) throw std::runtime_error(std::string{"pairwise_aligner: "} + pa_ctx_last_error(_ctx.get()));
    }

    void run(std::vector<uint8_t> const &b1, std::vector<uint64_t> const &o1, std::vector<uint8_t> const &b2,
             std::vector<uint64_t> const &o2, std::vector<int32_t> &scores, std::vector<uint8_t> &flags)
    {
        const uint64_t n = o1.size() - 1;
        scores.assign(n, 0);
        flags.assign(n, 0);
        const pa_config cfg = device_config();
        static const uint8_t empty = 0;
        check(pa_score_batch(_ctx.get(), &cfg, b1.empty() ? &empty : b1.data(), o1.data(),
                             b2.empty() ? &empty : b2.data(), o2.data(), n, nullptr, scores.data(), flags.data()));
    }

    // scores plus every pair's final DP column and row, whole cells: pa_score_batch_dp_cells,
    // sliced per pair
    void run_dp(std::vector<uint8_t> const &b1, std::vector<uint64_t> const &o1, std::vector<uint8_t> const &b2,
                std::vector<uint64_t> const &o2, std::vector<int32_t> &scores, std::vector<uint8_t> &flags,
                std::vector<std::vector<dp_cell>> &columns, std::vector<std::vector<dp_cell>> &rows)
    {
        const uint64_t n = o1.size() - 1;
        scores.assign(n, 0);
        flags.assign(n, 0);
        std::vector<int32_t> col(b1.size() + n + 1), row(b2.size() + n + 1), col_gap(col.size()),
            row_gap(row.size());
        const pa_config cfg = device_config();
        static const uint8_t empty = 0;
        check(pa_score_batch_dp_cells(_ctx.get(), &cfg, b1.empty() ? &empty : b1.data(), o1.data(),
                                      b2.empty() ? &empty : b2.data(), o2.data(), n, nullptr, scores.data(),
                                      flags.data(), col.data(), row.data(), col_gap.data(), row_gap.data()));
        columns.resize(n);
        rows.resize(n);
        for (uint64_t p = 0; p < n; ++p) {
            const uint64_t c0 = o1[p] + p, r0 = o2[p] + p;
            columns[p].clear();
            rows[p].clear();
            for (uint64_t i = 0; i <= o1[p + 1] - o1[p]; ++i) columns[p].emplace_back(col[c0 + i], col_gap[c0 + i]);
            for (uint64_t j = 0; j <= o2[p + 1] - o2[p]; ++j) rows[p].emplace_back(row[r0 + j], row_gap[r0 + j]);
        }
    }

    // one query against every target: pa_score_one_to_many (query profile on the device)
    void run_one_to_many(std::vector<uint8_t> const &query, std::vector<uint8_t> const &b2,
                         std::vector<uint64_t> const &o2, std::vector<int32_t> &scores, std::vector<uint8_t> &flags)
    {
        const uint64_t n = o2.size() - 1;
        scores.assign(n, 0);
        flags.assign(n, 0);
        const pa_config cfg = device_config();
        static const uint8_t empty = 0;
        check(pa_score_one_to_many(_ctx.get(), &cfg, query.empty() ? &empty : query.data(), query.size(),
                                   b2.empty() ? &empty : b2.data(), o2.data(), n, nullptr, scores.data(),
                                   flags.data()));
    }

public:
    aligner() = default;
    aligner(pa_config cfg, std::vector<int32_t> matrix, std::vector<uint8_t> symbols, int device = 0) :
        _cfg{cfg}, _matrix{std::move(matrix)}, _symbols{std::move(symbols)}, _device{device} {}

    pa_config const &config() const noexcept { return _cfg; }
    // Planner/launch parameters, e.g. from pa_tune (no counterpart in the reference).
    void tuning(pa_tuning const &t) { _tuning = t; }
    std::vector<int32_t> const &matrix() const noexcept { return _matrix; }
    std::vector<uint8_t> const &symbols() const noexcept { return _symbols; }

    // one pair (interface_one_to_one_single.hpp:40-67)
    template <detail::sequence sequence1_t, detail::sequence sequence2_t>
    auto compute(sequence1_t &&sequence1, sequence2_t &&sequence2)
    {
        std::vector<uint8_t> b1, b2;
        std::vector<uint64_t> o1{0}, o2{0};
        detail::append(b1, o1, sequence1);
        detail::append(b2, o2, sequence2);
        std::vector<int32_t> scores;
        std::vector<uint8_t> flags;
        run(b1, o1, b2, o2, scores, flags);
        return aligner_result<sequence1_t, sequence2_t>{std::forward<sequence1_t>(sequence1),
                                                        std::forward<sequence2_t>(sequence2), scores[0], flags[0]};
    }

    // one pair with the initial DP column and row (interface_one_to_one_single.hpp:52-67).  Like the
    // reference, which re-initialises both from the gap model and the end-gap rules before the
    // recursion (dp_vector_single.hpp:67-80, affine_dp_algorithm.hpp:48-72), the device takes the
    // boundary from the configuration; the result carries the cells of the final DP column and
    // row (aligner_result.hpp:41-59) through pa_score_batch_dp_cells.
    template <detail::sequence sequence1_t, detail::sequence sequence2_t, typename dp_column_t, typename dp_row_t>
    auto compute(sequence1_t &&sequence1, sequence2_t &&sequence2, dp_column_t &&, dp_row_t &&)
    {
        std::vector<uint8_t> b1, b2;
        std::vector<uint64_t> o1{0}, o2{0};
        detail::append(b1, o1, sequence1);
        detail::append(b2, o2, sequence2);
        std::vector<int32_t> scores;
        std::vector<uint8_t> flags;
        std::vector<std::vector<dp_cell>> cols, rows;
        run_dp(b1, o1, b2, o2, scores, flags, cols, rows);
        return aligner_result<sequence1_t, sequence2_t>{std::forward<sequence1_t>(sequence1),
                                                        std::forward<sequence2_t>(sequence2), scores[0], flags[0],
                                                        std::move(cols[0]), std::move(rows[0])};
    }

    // one-to-one bulk (interface_one_to_one_bulk.hpp:47-100), any bulk size
    template <detail::sequence_bulk bulk1_t, detail::sequence_bulk bulk2_t>
    auto compute(bulk1_t &&bulk1, bulk2_t &&bulk2)
    {
        std::vector<uint8_t> b1, b2;
        std::vector<uint64_t> o1{0}, o2{0};
        for (auto &&s : bulk1) detail::append(b1, o1, s);
        for (auto &&s : bulk2) detail::append(b2, o2, s);
        if (o1.size() != o2.size()) throw std::invalid_argument("bulk compute needs equally many sequences");
        using result_t = bulk_result<bulk1_t, bulk2_t>;
        auto shared = std::make_shared<result_t>(result_t{std::forward<bulk1_t>(bulk1), std::forward<bulk2_t>(bulk2),
                                                          {}, {}});
        run(b1, o1, b2, o2, shared->_scores, shared->_flags);
        std::vector<aligner_result_bulk<result_t, false>> results;
        results.reserve(o1.size() - 1);
        for (std::size_t i = 0; i + 1 < o1.size(); ++i) results.emplace_back(shared, i);
        return results;
    }

    // one-to-one bulk with the initial DP column and row (interface_one_to_one_bulk.hpp:73-100):
    // as for one pair, the boundary is the configuration's; every result carries its pair's
    // final DP column and row (dp_column() / dp_row())
    template <detail::sequence_bulk bulk1_t, detail::sequence_bulk bulk2_t, typename dp_column_t, typename dp_row_t>
    auto compute(bulk1_t &&bulk1, bulk2_t &&bulk2, dp_column_t &&, dp_row_t &&)
    {
        std::vector<uint8_t> b1, b2;
        std::vector<uint64_t> o1{0}, o2{0};
        for (auto &&s : bulk1) detail::append(b1, o1, s);
        for (auto &&s : bulk2) detail::append(b2, o2, s);
        if (o1.size() != o2.size()) throw std::invalid_argument("bulk compute needs equally many sequences");
        using result_t = bulk_result<bulk1_t, bulk2_t>;
        auto shared = std::make_shared<result_t>(result_t{std::forward<bulk1_t>(bulk1), std::forward<bulk2_t>(bulk2),
                                                          {}, {}});
        run_dp(b1, o1, b2, o2, shared->_scores, shared->_flags, shared->_dp_columns, shared->_dp_rows);
        std::vector<aligner_result_bulk<result_t, false>> results;
        results.reserve(o1.size() - 1);
        for (std::size_t i = 0; i + 1 < o1.size(); ++i) results.emplace_back(shared, i);
        return results;
    }

    // one-to-many with the initial DP column and row (interface_one_to_many_bulk.hpp:67-77): the
    // query scored against every target, each result with its pair's final DP column and row
    template <detail::sequence sequence1_t, detail::sequence_bulk bulk2_t, typename dp_column_t, typename dp_row_t>
    auto compute(sequence1_t &&sequence1, bulk2_t &&bulk2, dp_column_t &&, dp_row_t &&)
    {
        std::vector<uint8_t> q, b1, b2;
        std::vector<uint64_t> oq{0}, o1{0}, o2{0};
        detail::append(q, oq, sequence1);
        for (auto &&s : bulk2) {
            detail::append(b2, o2, s);
            b1.insert(b1.end(), q.begin(), q.end());
            o1.push_back(b1.size());
        }
        using result_t = bulk_result<sequence1_t, bulk2_t>;
        auto shared = std::make_shared<result_t>(result_t{std::forward<sequence1_t>(sequence1),
                                                          std::forward<bulk2_t>(bulk2), {}, {}});
        run_dp(b1, o1, b2, o2, shared->_scores, shared->_flags, shared->_dp_columns, shared->_dp_rows);
        std::vector<aligner_result_bulk<result_t, true>> results;
        results.reserve(o2.size() - 1);
        for (std::size_t i = 0; i + 1 < o2.size(); ++i) results.emplace_back(shared, i);
        return results;
    }

    // one-to-many (interface_one_to_many_bulk.hpp:47-92): the query is planned once and, on the
    // device, scored through a query profile (score_model_matrix_simd_1xN.hpp:63-183)
    template <detail::sequence sequence1_t, detail::sequence_bulk bulk2_t>
    auto compute(sequence1_t &&sequence1, bulk2_t &&bulk2)
    {
        std::vector<uint8_t> q, b2;
        std::vector<uint64_t> oq{0}, o2{0};
        detail::append(q, oq, sequence1);
        for (auto &&s : bulk2) detail::append(b2, o2, s);
        using result_t = bulk_result<sequence1_t, bulk2_t>;
        auto shared = std::make_shared<result_t>(result_t{std::forward<sequence1_t>(sequence1),
                                                          std::forward<bulk2_t>(bulk2), {}, {}});
        run_one_to_many(q, b2, o2, shared->_scores, shared->_flags);
        std::vector<aligner_result_bulk<result_t, true>> results;
        results.reserve(o2.size() - 1);
        for (std::size_t i = 0; i + 1 < o2.size(); ++i) results.emplace_back(shared, i);
        return results;
    }
};

namespace cfg {

namespace detail_configure {

template <typename rule_t>
void collect(rule_t const &r, pa_config &cfg, std::vector<int32_t> &matrix, std::vector<uint8_t> &symbols,
             bool &method_seen, bool &lanes_seen)
{
    using namespace cfg::detail;
    if constexpr (!std::is_same_v<rule_t, initial_rule>) {
        using payload_t = std::remove_cvref_t<decltype(r.payload)>;
        if constexpr (std::is_same_v<payload_t, unitary_payload>) {
            cfg.score_kind = PA_SCORE_UNITARY;
            cfg.match = r.payload.match;
            cfg.mismatch = r.payload.mismatch;
            if (!lanes_seen) cfg.lane_policy = r.payload.lanes;
        } else if constexpr (std::is_same_v<payload_t, gap_payload>) {
            cfg.gap_open = r.payload.open;
            cfg.gap_extension = r.payload.extension;
        } else if constexpr (std::is_same_v<payload_t, method_payload>) {
            method_seen = true;
            cfg.method = r.payload.local ? PA_METHOD_LOCAL : PA_METHOD_GLOBAL;
            cfg.first_column_free = r.payload.leading.first_column == end_gap::free;
            cfg.first_row_free = r.payload.leading.first_row == end_gap::free;
            cfg.last_column_free = r.payload.trailing.last_column == end_gap::free;
            cfg.last_row_free = r.payload.trailing.last_row == end_gap::free;
        } else if constexpr (std::is_same_v<payload_t, lanes_payload>) {
            lanes_seen = true;
            cfg.lane_policy = r.payload.policy;
        } else {   // matrix payload
            cfg.score_kind = PA_SCORE_MATRIX;
            const auto &m = r.payload.matrix;
            const std::size_t dim = m.size();
            matrix.clear();
            symbols.clear();
            for (std::size_t i = 0; i < dim; ++i) {
                symbols.push_back(static_cast<uint8_t>(m[i].first));
                if (m[i].second.size() != dim) throw std::invalid_argument("substitution matrix must be square");
                for (std::size_t j = 0; j < dim; ++j) matrix.push_back(static_cast<int32_t>(m[i].second[j]));
            }
            cfg.matrix_dim = static_cast<int32_t>(dim);
            if (!lanes_seen) cfg.lane_policy = r.payload.lanes;
        }
        collect(r.predecessor, cfg, matrix, symbols, method_seen, lanes_seen);
    }
}

} // namespace detail_configure

namespace detail_configure {
template <typename rule_t>
void collect_lanes(rule_t const &r, pa_config &cfg, bool &lanes_seen)
{
    if constexpr (!std::is_same_v<rule_t, detail::initial_rule>) {
        if constexpr (std::is_same_v<std::remove_cvref_t<decltype(r.payload)>, detail::lanes_payload>) {
            cfg.lane_policy = r.payload.policy;
            lanes_seen = true;
        }
        collect_lanes(r.predecessor, cfg, lanes_seen);
    }
}
} // namespace detail_configure

// configure_aligner.hpp:194-206
struct configure_aligner_fn {
    template <detail::rule_type rule_t>
    auto operator()(rule_t const &rule, int device = 0) const
    {
        using r = std::remove_cvref_t<rule_t>;
        static_assert(detail::count_rules<r>(detail::rule_category::score_model) == 1,
                      "configure_aligner needs exactly one score model rule");
        static_assert(detail::count_rules<r>(detail::rule_category::gap_model) == 1,
                      "configure_aligner needs exactly one gap model rule");
        static_assert(detail::count_rules<r>(detail::rule_category::method) <= 1,
                      "configure_aligner takes at most one method rule");
        pa_config cfg{};
        cfg.method = PA_METHOD_GLOBAL;   // default method: global, penalised end gaps (initial.hpp:49-65)
        cfg.lane_policy = PA_LANES_AUTO;
        std::vector<int32_t> matrix;
        std::vector<uint8_t> symbols;
        bool method_seen = false, lanes_seen = false;
        // the lane rule may sit anywhere in the chain; read it first so it wins over the score model's
        detail_configure::collect_lanes(rule, cfg, lanes_seen);
        detail_configure::collect(rule, cfg, matrix, symbols, method_seen, lanes_seen);
        return aligner{cfg, std::move(matrix), std::move(symbols), device};
    }
};


inline constexpr configure_aligner_fn configure_aligner{};

} // namespace cfg
} // inline namespace v1
} // namespace seqan::pairwise_aligner

// structured bindings of a cell, as for the reference's affine cell (affine/affine_cell.hpp:86-99)
template <>
struct std::tuple_size<seqan::pairwise_aligner::dp_cell> : std::integral_constant<std::size_t, 2> {};
template <std::size_t I>
struct std::tuple_element<I, seqan::pairwise_aligner::dp_cell> {
    using type = int32_t;
};
