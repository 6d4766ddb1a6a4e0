// simd_port.cpp -- CPU baseline of bench.py: the reference's inter-sequence SIMD scheme restated
// for x86 (AVX-512BW or AVX2, chosen at run time).  TEST INFRASTRUCTURE: only bench.py's
// cpu_baseline leg and tests/ load it; the device library never does.
//
// What it restates (reference files):
//   * one pair per SIMD lane, sequences transposed into lane-major vectors and padded at the
//     end (dp_vector_bulk.hpp:83-103), int16 lanes (32 per AVX-512 register, 16 per AVX2
//     register) or int32 lanes (half as many) -- simd_base.hpp:28-57;
//   * the Gotoh cell applied to whole vectors (affine/affine_dp_algorithm.hpp:139-156), the
//     boundary of affine_initialisation_strategy.hpp:38-51, local clamp (dp_matrix_local.hpp:39-47);
//   * scores: unitary equality (score_model_unitary_simd.hpp:51-58) or a substitution-matrix
//     gather per lane (score_model_matrix_simd_NxN.hpp; the reference gathers int16 entries one
//     lane at a time, simd_selector.hpp:35-55, this port uses the hardware gather);
//   * extraction: H(m,n) per lane with the free-end maxima (tracker_global_scalar.hpp:36-61), or
//     the running maximum (tracker_local_simd_fixed.hpp:38-47).
// Lane width: int16 where the bound of the padded bulk proves that no H or gap state leaves the
// int16 range (lower = 3go + ge(M+N), upper = smax*min(M,N)), int32 otherwise.  Pairs are sorted
// by (method, m, n) before they are cut into bulks, and threads take bulks from a shared counter.
// Pad rows/columns score -1, so with gap scores <= 0 they never exceed the real cells.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

extern "C" {
typedef struct pao_config {
    int32_t mode;               /* 0 global, 1 local */
    int32_t first_column_free;
    int32_t first_row_free;
    int32_t last_column_free;
    int32_t last_row_free;
    int32_t gap_open;
    int32_t gap_extension;
    int32_t score_kind;         /* 0 unitary, 1 matrix */
    int32_t match;
    int32_t mismatch;
    int32_t dim;
    const int32_t *matrix;
    const uint8_t *symbols;
} pao_config;
}

namespace {

struct Bulk {
    int64_t pair[64];
    int m[64], n[64];
    int M = 0, N = 0;
    const uint8_t *const *s1 = nullptr;   // per pair: seq1 / seq2 bytes
    const uint8_t *const *s2 = nullptr;
    const int32_t *code = nullptr;        // byte -> residue code (unitary: byte value; matrix: rank)
    int32_t pad_code = 0;
    bool matrix = false;
    int32_t k1 = 0;                       // matrix: rank count + 1 (the pad rank is k1 - 1)
    const int32_t *table32 = nullptr;     // matrix: k1 x k1 scores
    int32_t match = 0, mismatch = 0;
    int64_t go = 0, ge = 0;
    bool local = false;
    int fc = 0, fr = 0, lc = 0, lr = 0;
};

} // namespace

#pragma GCC push_options
#pragma GCC target("avx512f,avx512bw")
namespace isa_avx512 {
#define VBYTES 64
#include "simd_port_kernel.inc"
#undef VBYTES
}
#pragma GCC pop_options

#pragma GCC push_options
#pragma GCC target("avx2")
namespace isa_avx2 {
#define VBYTES 32
#include "simd_port_kernel.inc"
#undef VBYTES
}
#pragma GCC pop_options

namespace {

// Scores of pairs with an empty sequence (closed forms, as the oracle).
int32_t empty_score(const pao_config &c, bool local, int64_t m, int64_t n)
{
    if (local) return INT32_MIN;
    const int64_t go = c.gap_open, ge = c.gap_extension;
    auto h0row = [&](int64_t j) -> int64_t { return (j == 0 || c.first_row_free) ? 0 : go + ge * j; };
    auto h0col = [&](int64_t i) -> int64_t { return (i == 0 || c.first_column_free) ? 0 : go + ge * i; };
    int64_t s;
    if (m == 0) {
        s = h0row(n);
        if (c.last_row_free && n > 0) s = std::max<int64_t>({s, 0, h0row(1)});
        if (c.last_column_free) s = std::max(s, h0row(n));
    } else {
        s = h0col(m);
        if (c.last_column_free) s = std::max<int64_t>({s, 0, h0col(1)});
        if (c.last_row_free) s = std::max(s, h0col(m));
    }
    return int32_t(s);
}

} // namespace

extern "C" int pas_isa(void)
{
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512f")) return 512;
    if (__builtin_cpu_supports("avx2")) return 256;
    return 0;
}

/* Scores n_pairs pairs with the inter-sequence SIMD baseline on `threads` threads.  pair_mode
 * (nullable) overrides cfg->mode per pair.  Returns 0, or -1 when the CPU has neither AVX2 nor
 * AVX-512BW, -2 for a residue outside the matrix alphabet, -3 for a score range beyond int32. */
extern "C" int pas_score_batch(const pao_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                               const uint8_t *seq2, const uint64_t *off2, int64_t n_pairs, const uint8_t *pair_mode,
                               int threads, int32_t *scores)
{
    const int isa = pas_isa();
    if (isa == 0) return -1;
    const pao_config &c = *cfg;
    // residue codes and the matrix table
    int32_t code[256];
    std::vector<int32_t> table;
    int32_t k1 = 0, pad = 256, smax = std::max(c.match, c.mismatch);
    if (c.score_kind == 1) {
        for (int b = 0; b < 256; ++b) code[b] = -1;
        for (int r = 0; r < c.dim; ++r) code[c.symbols[r]] = r;
        k1 = c.dim + 1;
        pad = c.dim;
        table.assign(size_t(k1) * k1, -1);
        smax = INT32_MIN;
        for (int a = 0; a < c.dim; ++a)
            for (int b = 0; b < c.dim; ++b) {
                table[size_t(a) * k1 + b] = c.matrix[a * c.dim + b];
                smax = std::max(smax, c.matrix[a * c.dim + b]);
            }
        for (int64_t p = 0; p < n_pairs; ++p) {
            for (uint64_t i = off1[p]; i < off1[p + 1]; ++i) if (code[seq1[i]] < 0) return -2;
            for (uint64_t i = off2[p]; i < off2[p + 1]; ++i) if (code[seq2[i]] < 0) return -2;
        }
    } else {
        for (int b = 0; b < 256; ++b) code[b] = b;
    }
    std::vector<const uint8_t *> p1(n_pairs), p2(n_pairs);
    std::vector<int64_t> order;
    order.reserve(n_pairs);
    for (int64_t p = 0; p < n_pairs; ++p) {
        p1[p] = seq1 + off1[p];
        p2[p] = seq2 + off2[p];
        const int64_t m = int64_t(off1[p + 1] - off1[p]), n = int64_t(off2[p + 1] - off2[p]);
        const bool local = pair_mode ? pair_mode[p] == 1 : c.mode == 1;
        if (m == 0 || n == 0) scores[p] = empty_score(c, local, m, n);
        else order.push_back(p);
    }
    auto mode_of = [&](int64_t p) { return pair_mode ? int(pair_mode[p] == 1) : int(c.mode == 1); };
    auto len1 = [&](int64_t p) { return int64_t(off1[p + 1] - off1[p]); };
    auto len2 = [&](int64_t p) { return int64_t(off2[p + 1] - off2[p]); };
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        return std::make_tuple(mode_of(x), len1(x), len2(x), x) < std::make_tuple(mode_of(y), len1(y), len2(y), y);
    });
    // bulks of the int16 lane count; a bulk that needs int32 lanes is run as two halves
    const int L16 = isa == 512 ? 32 : 16;
    struct Range { size_t begin, end; };
    std::vector<Range> bulks;
    for (size_t i = 0; i < order.size();) {
        size_t e = i + 1;
        while (e < order.size() && e - i < size_t(L16) && mode_of(order[e]) == mode_of(order[i])) ++e;
        bulks.push_back({i, e});
        i = e;
    }
    std::atomic<size_t> next{0};
    std::atomic<int> failed{0};
    auto worker = [&]() {
        std::vector<int16_t> A16, B16;
        std::vector<int32_t> A32, B32;
        std::vector<uint8_t> work;
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= bulks.size()) return;
            const Range r = bulks[k];
            Bulk bk;
            bk.s1 = p1.data();
            bk.s2 = p2.data();
            bk.code = code;
            bk.pad_code = pad;
            bk.matrix = c.score_kind == 1;
            bk.k1 = k1;
            bk.table32 = table.data();
            bk.match = c.match;
            bk.mismatch = c.mismatch;
            bk.go = c.gap_open;
            bk.ge = c.gap_extension;
            bk.local = mode_of(order[r.begin]) == 1;
            bk.fc = c.first_column_free; bk.fr = c.first_row_free;
            bk.lc = c.last_column_free; bk.lr = c.last_row_free;
            int M = 0, N = 0;
            for (size_t i = r.begin; i < r.end; ++i) {
                M = std::max<int>(M, int(len1(order[i])));
                N = std::max<int>(N, int(len2(order[i])));
            }
            const int64_t lower = bk.local ? bk.go + bk.ge : 3 * bk.go + bk.ge * (int64_t(M) + N);
            const int64_t upper = int64_t(std::max(smax, 0)) * std::min(M, N) + 1;
            if (lower < INT32_MIN / 4 || upper > INT32_MAX / 4) { failed = -3; return; }
            const bool w16 = lower > -32700 && upper < 32767;
            const int L = w16 ? L16 : L16 / 2;
            for (size_t h = r.begin; h < r.end; h += size_t(L)) {
                const size_t he = std::min(r.end, h + size_t(L));
                bk.M = 0;
                bk.N = 0;
                for (int l = 0; l < 64; ++l) bk.pair[l] = -1;
                for (size_t i = h; i < he; ++i) {
                    const int l = int(i - h);
                    bk.pair[l] = order[i];
                    bk.m[l] = int(len1(order[i]));
                    bk.n[l] = int(len2(order[i]));
                    bk.M = std::max(bk.M, bk.m[l]);
                    bk.N = std::max(bk.N, bk.n[l]);
                }
                if (isa == 512) {
                    if (w16) isa_avx512::run_bulk<int16_t>(bk, A16, B16, work, scores);
                    else isa_avx512::run_bulk<int32_t>(bk, A32, B32, work, scores);
                } else {
                    if (w16) isa_avx2::run_bulk<int16_t>(bk, A16, B16, work, scores);
                    else isa_avx2::run_bulk<int32_t>(bk, A32, B32, work, scores);
                }
            }
        }
    };
    const int nt = std::max(1, threads);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
    return failed.load();
}
