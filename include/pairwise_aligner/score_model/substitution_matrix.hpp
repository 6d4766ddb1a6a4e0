// BLOSUM62 substitution tables in the two symbol orders the reference ships
// (include/pairwise_aligner/score_model/substitution_matrix.hpp:28-87 of rrahn/pairwise_aligner):
//   blosum62_standard<T>  21 symbols "ACDEFGHIKLMNPQRSTVWXY"
//   blosum62_extended<T>  27 symbols "ABCDEFGHIJKLMNOPQRSTUVWXYZ*"
// The values are the public NCBI BLOSUM62 matrix, stored once below as NCBI-format text and
// re-ordered at compile time.  The extended order adds J (row given separately, with the
// values the reference uses) and O and U, which score exactly like X.
// The same text block is read by the Python mirror (pairwise_aligner_amd/tables.py).
#pragma once

#include <array>
#include <cstdint>
#include <string_view>
#include <utility>

namespace seqan::pairwise_aligner {
inline namespace v1 {
namespace detail {

// NCBI BLOSUM62, symbol order of the header line.
inline constexpr std::string_view blosum62_ncbi_text = R"BLOSUM(
   A  R  N  D  C  Q  E  G  H  I  L  K  M  F  P  S  T  W  Y  V  B  Z  X  *
A  4 -1 -2 -2  0 -1 -1  0 -2 -1 -1 -1 -1 -2 -1  1  0 -3 -2  0 -2 -1  0 -4
R -1  5  0 -2 -3  1  0 -2  0 -3 -2  2 -1 -3 -2 -1 -1 -3 -2 -3 -1  0 -1 -4
N -2  0  6  1 -3  0  0  0  1 -3 -3  0 -2 -3 -2  1  0 -4 -2 -3  3  0 -1 -4
D -2 -2  1  6 -3  0  2 -1 -1 -3 -4 -1 -3 -3 -1  0 -1 -4 -3 -3  4  1 -1 -4
C  0 -3 -3 -3  9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1 -3 -3 -2 -4
Q -1  1  0  0 -3  5  2 -2  0 -3 -2  1  0 -3 -1  0 -1 -2 -1 -2  0  3 -1 -4
E -1  0  0  2 -4  2  5 -2  0 -3 -3  1 -2 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4
G  0 -2  0 -1 -3 -2 -2  6 -2 -4 -4 -2 -3 -3 -2  0 -2 -2 -3 -3 -1 -2 -1 -4
H -2  0  1 -1 -3  0  0 -2  8 -3 -3 -1 -2 -1 -2 -1 -2 -2  2 -3  0  0 -1 -4
I -1 -3 -3 -3 -1 -3 -3 -4 -3  4  2 -3  1  0 -3 -2 -1 -3 -1  3 -3 -3 -1 -4
L -1 -2 -3 -4 -1 -2 -3 -4 -3  2  4 -2  2  0 -3 -2 -1 -2 -1  1 -4 -3 -1 -4
K -1  2  0 -1 -3  1  1 -2 -1 -3 -2  5 -1 -3 -1  0 -1 -3 -2 -2  0  1 -1 -4
M -1 -1 -2 -3 -1  0 -2 -3 -2  1  2 -1  5  0 -2 -1 -1 -1 -1  1 -3 -1 -1 -4
F -2 -3 -3 -3 -2 -3 -3 -3 -1  0  0 -3  0  6 -4 -2 -2  1  3 -1 -3 -3 -1 -4
P -1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4  7 -1 -1 -4 -3 -2 -2 -1 -2 -4
S  1 -1  1  0 -1  0  0  0 -1 -2 -2  0 -1 -2 -1  4  1 -3 -2 -2  0  0  0 -4
T  0 -1  0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1  1  5 -2 -2  0 -1 -1  0 -4
W -3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1  1 -4 -3 -2 11  2 -3 -4 -3 -2 -4
Y -2 -2 -2 -3 -2 -1 -2 -3  2 -1 -1 -2 -1  3 -3 -2 -2  2  7 -1 -3 -2 -1 -4
V  0 -3 -3 -3 -1 -2 -2 -3 -3  3  1 -2  1 -1 -2 -2  0 -3 -1  4 -3 -2 -1 -4
B -2 -1  3  4 -3  0  1 -1  0 -3 -4  0 -3 -3 -2  0 -1 -4 -3 -3  4  1 -1 -4
Z -1  0  0  1 -3  3  4 -2  0 -3 -3  1 -1 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4
X  0 -1 -1 -1 -2 -1 -1 -1 -1 -1 -1 -1 -1 -1 -2  0  0 -2 -1 -1 -1 -1 -1 -4
* -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4  1
)BLOSUM";

// Row of J (I or L) against the header symbols above, then J against J.
inline constexpr std::string_view blosum62_j_text = R"BLOSUM(
J -1 -3 -3 -4 -1 -3 -3 -4 -3  3  3 -3  2  0 -3 -2 -1 -3 -1  2 -4 -3 -1 -4  3
)BLOSUM";

inline constexpr std::string_view blosum62_standard_order = "ACDEFGHIKLMNPQRSTVWXY";
inline constexpr std::string_view blosum62_extended_order = "ABCDEFGHIJKLMNOPQRSTUVWXYZ*";

struct ncbi_table {
    char symbols[32]{};
    int count{0};
    int values[32][32]{};
    int j_row[32]{};
    int j_self{0};
};

constexpr bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; }

constexpr int parse_int(std::string_view text, std::size_t & pos)
{
    while (is_space(text[pos])) ++pos;
    bool neg = false;
    if (text[pos] == '-') { neg = true; ++pos; }
    int v = 0;
    while (pos < text.size() && text[pos] >= '0' && text[pos] <= '9') v = v * 10 + (text[pos++] - '0');
    return neg ? -v : v;
}

constexpr ncbi_table parse_ncbi()
{
    ncbi_table t{};
    std::string_view text = blosum62_ncbi_text;
    std::size_t pos = 0;
    // header line
    while (is_space(text[pos])) ++pos;
    while (text[pos] != '\n') {
        if (!is_space(text[pos])) t.symbols[t.count++] = text[pos];
        ++pos;
    }
    for (int r = 0; r < t.count; ++r) {
        while (is_space(text[pos])) ++pos;
        ++pos;  // row symbol
        for (int c = 0; c < t.count; ++c) t.values[r][c] = parse_int(text, pos);
    }
    std::string_view jt = blosum62_j_text;
    std::size_t jp = 0;
    while (is_space(jt[jp])) ++jp;
    ++jp;  // 'J'
    for (int c = 0; c < t.count; ++c) t.j_row[c] = parse_int(jt, jp);
    t.j_self = parse_int(jt, jp);
    return t;
}

inline constexpr ncbi_table blosum62_ncbi = parse_ncbi();

constexpr int ncbi_index(char s)
{
    if (s == 'O' || s == 'U') s = 'X';
    for (int i = 0; i < blosum62_ncbi.count; ++i)
        if (blosum62_ncbi.symbols[i] == s) return i;
    return -1;  // 'J'
}

constexpr int blosum62_value(char a, char b)
{
    int ia = ncbi_index(a), ib = ncbi_index(b);
    if (ia < 0 && ib < 0) return blosum62_ncbi.j_self;
    if (ia < 0) return blosum62_ncbi.j_row[ib];
    if (ib < 0) return blosum62_ncbi.j_row[ia];
    return blosum62_ncbi.values[ia][ib];
}

template <typename score_t, std::size_t N>
constexpr std::array<std::pair<char, std::array<score_t, N>>, N> make_table(std::string_view order)
{
    std::array<std::pair<char, std::array<score_t, N>>, N> table{};
    for (std::size_t r = 0; r < N; ++r) {
        table[r].first = order[r];
        for (std::size_t c = 0; c < N; ++c)
            table[r].second[c] = static_cast<score_t>(blosum62_value(order[r], order[c]));
    }
    return table;
}

} // namespace detail

template <typename score_t = int32_t>
inline constexpr std::array<std::pair<char, std::array<score_t, 21>>, 21> blosum62_standard =
    detail::make_table<score_t, 21>(detail::blosum62_standard_order);

template <typename score_t = int32_t>
inline constexpr std::array<std::pair<char, std::array<score_t, 27>>, 27> blosum62_extended =
    detail::make_table<score_t, 27>(detail::blosum62_extended_order);

} // inline namespace v1
} // namespace seqan::pairwise_aligner
