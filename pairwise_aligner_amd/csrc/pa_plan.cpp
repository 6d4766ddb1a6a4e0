// pa_plan.cpp -- host-side planning for the striped DP kernels.  See pa_plan.hpp.
#include "pa_plan.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <immintrin.h>
#include <numeric>
#include <thread>
#include <tuple>

namespace pa {

bool VariantKey::operator<(const VariantKey &o) const
{
    return std::tie(width, G, R, score, local, track, rowmax, export_dp) <
           std::tie(o.width, o.G, o.R, o.score, o.local, o.track, o.rowmax, o.export_dp);
}
bool VariantKey::operator==(const VariantKey &o) const
{
    return std::tie(width, G, R, score, local, track, rowmax, export_dp) ==
           std::tie(o.width, o.G, o.R, o.score, o.local, o.track, o.rowmax, o.export_dp);
}

// Rows per lane that kernels are instantiated for (lanes per group G = 16).
static const uint8_t kRowsPerLane[] = {4, 6, 8, 10, 12, 14, 16};

Tuning resolve_tuning(const pa_tuning *t)
{
    Tuning r;
    if (!t) return r;
    if (t->cell_cost > 0) r.cell_cost = t->cell_cost;
    if (t->step_cost > 0) r.step_cost = t->step_cost;
    if (t->rows_mask & 0x7F) r.rows_mask = t->rows_mask & 0x7F;
    if (t->wavefront_min_rows) r.wavefront_min_rows = std::max<uint32_t>(t->wavefront_min_rows, 256);
    if (t->wavefront_tail > 0) r.wavefront_tail = t->wavefront_tail;
    else if (t->wavefront_tail < 0) r.wavefront_tail = 0;
    if (t->streams) r.streams = std::min<uint32_t>(t->streams, 4);
    r.eager = t->eager != 0;
    r.cost_order = t->cost_order != 0;
    r.persistent = std::min<uint32_t>(t->persistent, 4);
    return r;
}

void choose_shape(uint32_t m, const Tuning &tuning, bool wavefront, uint8_t &G, uint8_t &R)
{
    if (wavefront) {   // long pair: one workgroup of four waves, stripes of 64*16 rows
        G = 64;
        R = 16;
        return;
    }
    // Modelled cost S * (n + G - 1) * (cell_cost * R + step_cost) for S stripes of G*R rows over
    // n columns; n + G - 1 is the same for every R, so the best R depends on m alone.
    G = 16;
    double best = 1e300;
    R = 16;
    for (int i = 0; i < 7; ++i) {
        if (!(tuning.rows_mask & (1u << i))) continue;
        const uint8_t r = kRowsPerLane[i];
        const uint32_t rows = uint32_t(G) * r;
        const uint32_t S = (m + rows - 1) / rows;
        const double cost = double(S) * (tuning.cell_cost * r + tuning.step_cost);
        if (cost < best * 0.999) { best = cost; R = r; }
    }
}

double striped_cost(uint32_t m, uint32_t n, const Tuning &tuning)
{
    uint8_t G, R;
    choose_shape(m, tuning, false, G, R);
    const uint32_t rows = uint32_t(G) * R;
    return double((m + rows - 1) / rows) * (double(n) + G - 1) * (tuning.cell_cost * R + tuning.step_cost);
}

RangeBound range_bound(const Plan &plan, bool local, uint32_t m, uint32_t n)
{
    const int64_t go = std::min<int64_t>(plan.go, 0), ge = std::min<int64_t>(plan.ge, 0);
    const int64_t sp = std::max<int64_t>(plan.smax, 0);
    RangeBound b;
    // + the largest drift the kernels add: |ge| * (R + 1) along the rows (R <= 16) and
    // |ge| * G along the columns (G <= 64)
    b.upper = sp * int64_t(std::min(m, n)) - ge * 81;
    if (local)
        b.lower = go + ge;                       // H >= 0, gap states >= H + go + ge
    else   // L-shaped path plus one gap opening, and the diagonal value kept one column back
        b.lower = 3 * go + ge * (int64_t(m) + int64_t(n) + 1);
    return b;
}

namespace {

unsigned planner_threads()
{
    return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// Runs body(i) for i in [0, n) on up to 16 threads, in blocks of `grain` indices.
template <typename F>
void parallel_for(size_t n, F &&body, size_t grain = 256)
{
    const unsigned nt = unsigned(std::min<size_t>(planner_threads(), (n + grain - 1) / grain));
    if (nt <= 1) {
        for (size_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const size_t b = next.fetch_add(grain);
            if (b >= n) return;
            for (size_t i = b; i < std::min(n, b + grain); ++i) body(i);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned k = 1; k < nt; ++k) pool.emplace_back(worker);
    worker();
    for (auto &th : pool) th.join();
}

// Sorts every vector of `vs` by cmp on up to 16 threads: each vector is cut into up to 16 runs
// (one per 16Ki elements, a power of two), all runs of all vectors are sorted concurrently and
// then merged pairwise.  For a strict total order (or with `stable`) the result is the one
// std::sort (std::stable_sort) gives.
template <typename T, typename Cmp>
void parallel_sort(std::vector<std::vector<T>> &vs, Cmp cmp, bool stable)
{
    struct Run { size_t v, lo, hi; };
    std::vector<std::vector<size_t>> cuts(vs.size());
    std::vector<Run> runs;
    for (size_t v = 0; v < vs.size(); ++v) {
        const size_t n = vs[v].size();
        size_t parts = 1;
        while (parts < 16 && n / (parts * 2) >= 16384) parts *= 2;
        for (size_t i = 0; i <= parts; ++i) cuts[v].push_back(n * i / parts);
        for (size_t i = 0; i < parts; ++i) runs.push_back(Run{v, cuts[v][i], cuts[v][i + 1]});
    }
    parallel_for(runs.size(), [&](size_t r) {
        auto first = vs[runs[r].v].begin();
        if (stable) std::stable_sort(first + runs[r].lo, first + runs[r].hi, cmp);
        else std::sort(first + runs[r].lo, first + runs[r].hi, cmp);
    }, 1);
    for (size_t w = 1; w < 16; w *= 2) {
        std::vector<std::pair<size_t, size_t>> merges;   // (vector, first run)
        for (size_t v = 0; v < vs.size(); ++v)
            for (size_t i = 0; i + w < cuts[v].size() - 1; i += 2 * w) merges.push_back({v, i});
        parallel_for(merges.size(), [&](size_t k) {
            const auto [v, i] = merges[k];
            const std::vector<size_t> &c = cuts[v];
            auto first = vs[v].begin();
            std::inplace_merge(first + c[i], first + c[i + w], first + c[std::min(i + 2 * w, c.size() - 1)], cmp);
        }, 1);
    }
}

struct Alphabet {
    bool dna = false;
    int16_t code[256];   // symbol -> code (DNA 0..3) or rank (matrix), -1 if absent
    uint32_t size = 0;
};

// Marks the byte values in [p, p + n).  Bytes already marked are recognised 32 at a time by a
// nibble-table membership test (AVX2 byte shuffles: lo[l] has bit h set when byte h*16 + l is
// marked, for h < 8; hi[h] = 1 << h), the rest one at a time.
__attribute__((target("avx2"))) void mark_bytes_avx2(const uint8_t *p, uint64_t n, std::array<uint8_t, 256> &m)
{
    alignas(32) uint8_t lo[32], hi[32];
    auto tables = [&]() {
        for (int l = 0; l < 16; ++l) {
            uint8_t bits = 0;
            for (int h = 0; h < 8; ++h) bits |= m[h * 16 + l] ? uint8_t(1u << h) : 0;
            lo[l] = lo[l + 16] = bits;
            hi[l] = hi[l + 16] = l < 8 ? uint8_t(1u << l) : 0;
        }
    };
    tables();
    __m256i vlo = _mm256_load_si256(reinterpret_cast<const __m256i *>(lo));
    __m256i vhi = _mm256_load_si256(reinterpret_cast<const __m256i *>(hi));
    const __m256i nib = _mm256_set1_epi8(0x0F), zero = _mm256_setzero_si256();
    uint64_t i = 0;
    for (; i + 32 <= n; i += 32) {
        const __m256i x = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(p + i));
        const __m256i a = _mm256_shuffle_epi8(vlo, _mm256_and_si256(x, nib));
        const __m256i b = _mm256_shuffle_epi8(vhi, _mm256_and_si256(_mm256_srli_epi16(x, 4), nib));
        uint32_t miss = uint32_t(_mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_and_si256(a, b), zero)));
        if (!miss) continue;
        bool added = false;
        for (; miss; miss &= miss - 1) {
            const uint8_t c = p[i + uint32_t(__builtin_ctz(miss))];
            added |= !m[c];
            m[c] = 1;
        }
        if (added) {
            tables();
            vlo = _mm256_load_si256(reinterpret_cast<const __m256i *>(lo));
            vhi = _mm256_load_si256(reinterpret_cast<const __m256i *>(hi));
        }
    }
    for (; i < n; ++i) m[p[i]] = 1;
}

// Marks the byte values that occur in the batch's sequences (in parallel over blocks of bytes).
void scan_residues(const BatchView &b, bool (&seen)[256])
{
    const uint8_t *src[2] = {b.seq1 + b.off1[0], b.seq2 + b.off2[0]};
    const uint64_t len[2] = {b.end1() - b.off1[0], b.off2[b.n_pairs] - b.off2[0]};
    constexpr uint64_t kBlock = uint64_t(1) << 18;
    const uint64_t nb0 = (len[0] + kBlock - 1) / kBlock, nb1 = (len[1] + kBlock - 1) / kBlock;
    const bool avx2 = __builtin_cpu_supports("avx2");
    std::vector<std::array<uint8_t, 256>> marks(nb0 + nb1);
    parallel_for(nb0 + nb1, [&](size_t k) {
        const int side = k < nb0 ? 0 : 1;
        const uint64_t lo = (side ? k - nb0 : k) * kBlock, hi = std::min(len[side], lo + kBlock);
        std::array<uint8_t, 256> &m = marks[k];
        m.fill(0);
        if (avx2) {
            mark_bytes_avx2(src[side] + lo, hi - lo, m);
            return;
        }
        for (uint64_t i = lo; i < hi; ++i) m[src[side][i]] = 1;
    }, 1);
    for (const auto &m : marks)
        for (int c = 0; c < 256; ++c) seen[c] = seen[c] || m[c];
}

pa_status build_alphabet(const pa_config &cfg, const BatchView &b, Plan &plan, Alphabet &alpha, std::string &err)
{
    for (int c = 0; c < 256; ++c) alpha.code[c] = -1;
    if (cfg.score_kind == PA_SCORE_UNITARY) {
        bool seen[256] = {false};
        if (b.known_codes)
            for (int c = 0; c < 256; ++c) seen[c] = b.known_codes[c] >= 0;
        else
            scan_residues(b, seen);
        uint32_t k = 0;
        for (int c = 0; c < 256; ++c) if (seen[c]) alpha.code[c] = int16_t(k++);
        alpha.size = k;
        if (cfg.match < -32768 || cfg.match > 32767 || cfg.mismatch < -32768 || cfg.mismatch > 32767) {
            err = "match/mismatch scores must fit in 16 bits";
            return PA_ERR_UNSUPPORTED;
        }
        plan.smax = std::max(cfg.match, cfg.mismatch);
        plan.smin = std::min(cfg.match, cfg.mismatch);
        // global-method kernels add their drift (up to -2ge) to every substitution score
        const int32_t shift_max = -(kScoreShifts - 1) * cfg.gap_extension;
        if (cfg.match + shift_max > 32767 || cfg.mismatch + shift_max > 32767) {
            err = "match/mismatch minus twice gap_extension must fit in 16 bits";
            return PA_ERR_UNSUPPORTED;
        }
        if (k <= 4) {
            alpha.dna = true;
            plan.score = kScoreDna;
            for (int v = 0; v < kScoreShifts; ++v) {
                const uint32_t M = uint32_t(cfg.match - v * cfg.gap_extension) & 0xFFFFu;
                const uint32_t X = uint32_t(cfg.mismatch - v * cfg.gap_extension) & 0xFFFFu;
                plan.tbl_lo[v] = M | (X << 16);
                plan.tbl_hi[v] = X | (X << 16);
            }
            return PA_OK;
        }
        if (k > kMaxMatrixSymbols) {
            err = "unitary scoring over more than 31 distinct symbols is not supported on the device";
            return PA_ERR_UNSUPPORTED;
        }
        plan.score = kScoreMatrix;
        plan.alphabet = k + 1;
        plan.stride = (plan.alphabet + 1) & ~1u;
        for (int v = 0; v < kScoreShifts; ++v) {
            std::vector<int16_t> &t = plan.table[v];
            t.assign(size_t(plan.alphabet) * plan.stride, int16_t(-1));
            for (uint32_t a = 0; a < k; ++a)
                for (uint32_t c = 0; c < k; ++c)
                    t[a * plan.stride + c] = int16_t((a == c ? cfg.match : cfg.mismatch) - v * cfg.gap_extension);
        }
        return PA_OK;
    }
    if (cfg.score_kind != PA_SCORE_MATRIX || cfg.matrix == nullptr || cfg.matrix_symbols == nullptr) {
        err = "invalid score model";
        return PA_ERR_INVALID_ARGUMENT;
    }
    const int32_t K = cfg.matrix_dim;
    if (K < 1 || K > int32_t(kMaxMatrixSymbols)) {
        err = "substitution matrix dimension must be 1..31";
        return PA_ERR_UNSUPPORTED;
    }
    for (int32_t r = 0; r < K; ++r) alpha.code[cfg.matrix_symbols[r]] = int16_t(r);
    alpha.size = uint32_t(K);
    plan.score = kScoreMatrix;
    plan.alphabet = uint32_t(K) + 1;
    plan.stride = (plan.alphabet + 1) & ~1u;
    for (std::vector<int16_t> &t : plan.table) t.assign(size_t(plan.alphabet) * plan.stride, int16_t(-1));
    plan.smax = INT32_MIN;
    plan.smin = INT32_MAX;
    for (int32_t a = 0; a < K; ++a)
        for (int32_t c = 0; c < K; ++c) {
            int32_t v = cfg.matrix[a * K + c];
            // global-method kernels add their drift (up to -2ge) to every substitution score
            if (v < -32768 || v - (kScoreShifts - 1) * cfg.gap_extension > 32767) {
                err = "substitution scores (minus twice gap_extension) must fit in 16 bits";
                return PA_ERR_UNSUPPORTED;
            }
            for (int s = 0; s < kScoreShifts; ++s)
                plan.table[s][a * plan.stride + c] = int16_t(v - s * cfg.gap_extension);
            plan.smax = std::max(plan.smax, v);
            plan.smin = std::min(plan.smin, v);
        }
    // every residue must be in the alphabet (alphabet_rank_map_scalar.hpp:59-63 asserts it)
    if (b.known_codes) return PA_OK;
    bool seen[256] = {false};
    scan_residues(b, seen);
    for (int c = 0; c < 256; ++c)
        if (seen[c] && alpha.code[c] < 0) { err = "residue outside the substitution matrix alphabet"; return PA_ERR_INVALID_SYMBOL; }
    return PA_OK;
}

struct Item {
    uint32_t pair;
    uint32_t m, n;
    uint32_t S;
};

} // namespace

double task_cost(const VariantKey &key, const Task &t, const Tuning &tuning)
{
    return double(t.S) * (double(t.N) + key.G - 1) * (tuning.cell_cost * key.R + tuning.step_cost) *
           (key.G == 64 ? 4.0 : 1.0);
}

uint32_t lds_bytes(const Plan &plan, const VariantKey &key)
{
    if (key.score == kScoreProfile) return plan.profile_ranks * plan.profile_rows * 2;
    if (key.score == kScorePairs) return pair_table_bytes(plan.alphabet);
    if (key.score == kScoreMatrix) return lds_table_bytes(plan.alphabet, plan.stride);
    return 0;
}

void code_tables(const Plan &plan, const VariantKey &key, CodeTables &out)
{
    const bool w16 = key.width == 16;
    const uint32_t pad = plan.pad_rank;
    for (int sym = 0; sym < 256; ++sym) {
        const int16_t code = plan.code[sym];
        const uint32_t c = code < 0 ? 0u : uint32_t(code);   // absent bytes never occur
        uint32_t r = 0, k = 0;
        switch (key.score) {
        case kScoreDna:
            r = w16 ? dna_row16(c) : dna_row32(c);
            k = w16 ? dna_col16(c) : dna_col32(c);
            break;
        case kScorePairs: r = c; k = pair_col_offset(c); break;   // rows: the rank itself
        case kScoreProfile: r = 0; k = profile_col_offset(c, plan.profile_rows); break;
        default: r = lds_row_offset(c, plan.stride); k = lds_col_offset(c); break;
        }
        out.row[sym] = r;
        out.col[sym] = k;
    }
    switch (key.score) {
    case kScoreDna:
        out.row_pad = w16 ? kDnaRowPad16 : kDnaRowPad32;
        out.col_pad = w16 ? kDnaColPad16 : kDnaColPad32;
        break;
    case kScorePairs: out.row_pad = pad; out.col_pad = pair_col_offset(pad); break;
    case kScoreProfile: out.row_pad = 0; out.col_pad = profile_col_offset(plan.profile_ranks - 1, plan.profile_rows); break;
    default: out.row_pad = lds_row_offset(pad, plan.stride); out.col_pad = lds_col_offset(pad); break;
    }
}

// Fused launches: the striped int16 local launches whose keys differ only in R become one
// persistent launch with one work queue, whose wave items (runs of up to four consecutive tasks of
// one member launch, alike in shape) come in decreasing order of modelled time across the members
// -- one longest-first order for the whole group instead of one per launch, and no launch of the
// group waiting for a stream or forming the pass's tail alone.  The kernel that holds every R runs
// at three waves per SIMD (167 VGPRs with the pair-row table, 150 with the byte-permute scores),
// the occupancy of the pair-row-table sweep at R = 16 alone; the byte-permute sweeps alone would
// hold four, and fused they still run faster (the mixed stream's DNA half at 100,000 pairs: 84.2
// -> 79.9 ms; the uneven 1M batch: 38.0 -> 35.1 ms).  Fused unless Tuning::persistent is 1 or 3.
void fuse_launches(Plan &plan, const float *pair_cost)
{
    const size_t nl = plan.launches.size();
    plan.fused.clear();
    plan.fused_items.clear();
    plan.fused_of.assign(nl, -1);
    const uint32_t form = plan.tuning.persistent;
    if (form == 1 || form == 3) return;
    // the global method's launches are fused in form 4 only
    auto fuses = [&](const VariantKey &k) { return fused_launchable(k) && (k.local || form == 4); };
    // launches that differ only in R, within one range of the fused kernel (fused_rows)
    auto group_key = [](VariantKey k) { k.R = uint8_t(fused_rows(k).lo); return k; };
    // modelled work of every launch, local cells weighed 10/8 (ops per cell)
    std::vector<double> work(nl, 0.0);
    double total = 0;
    for (size_t i = 0; i < nl; ++i) {
        const Launch &L = plan.launches[i];
        for (uint32_t t = L.task_begin; t < L.task_end; ++t) work[i] += task_cost(L.key, plan.tasks[t], plan.tuning);
        if (L.key.local) work[i] *= 1.25;
        total += work[i];
    }
    std::vector<bool> taken(nl, false);
    for (size_t i = 0; i < nl; ++i) {
        if (taken[i] || !fuses(plan.launches[i].key)) continue;
        const VariantKey gk = group_key(plan.launches[i].key);
        std::vector<uint32_t> members;
        for (size_t j = i; j < nl; ++j) {
            const Launch &L = plan.launches[j];
            if (taken[j] || !fuses(L.key) || !(group_key(L.key) == gk)) continue;
            if (L.task_end > kFusedMaxTasks) continue;   // beyond the item's task field
            members.push_back(uint32_t(j));
            taken[j] = true;
        }
        if (members.size() < 2) continue;
        struct Item { double time; uint32_t item; };
        std::vector<Item> items;
        for (uint32_t li : members) {
            const Launch &L = plan.launches[li];
            for (uint32_t k0 = L.task_begin; k0 < L.task_end; k0 += 4) {
                const uint32_t n = std::min<uint32_t>(4, L.task_end - k0);
                double t = 0;   // modelled time, or with pair_cost the predicted cost of the costliest pair
                for (uint32_t k = k0; k < k0 + n; ++k) {
                    if (!pair_cost) {
                        t = std::max(t, task_cost(L.key, plan.tasks[k], plan.tuning));
                        continue;
                    }
                    for (int h = 0; h < 2; ++h)
                        if (plan.tasks[k].pair[h] != kNone) t = std::max(t, double(pair_cost[plan.tasks[k].pair[h]]));
                }
                items.push_back(Item{t, fused_item(L.key.R, n, k0)});
            }
        }
        std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.time > y.time; });
        FusedLaunch f;
        f.leader = members[0];
        double w = 0;
        for (uint32_t li : members) w += work[li];
        f.share = total > 0 ? w / total : 1.0;
        f.items_begin = uint32_t(plan.fused_items.size());
        for (const Item &it : items) plan.fused_items.push_back(it.item);
        f.items_end = uint32_t(plan.fused_items.size());
        for (uint32_t li : members) plan.fused_of[li] = int32_t(plan.fused.size());
        plan.fused.push_back(f);
    }
}

void order_launches(Plan &plan)
{
    // Wavefront launches first: each of their workgroups is one long serial sweep, so they
    // set the critical path (longest task first).  Then the striped launches by the time each
    // would take on its own: its longest task, or its whole work spread over the groups the GPU
    // holds at once, whichever is larger -- so that neither a launch of a few long tasks nor a
    // large launch of short ones ends up in the tail of the pass.  Streams round robin.  A fused
    // launch counts once, with the tasks of all its members, in its leader's place; the other
    // members come last and share the leader's stream (they are not launched on their own).
    const size_t nl = plan.launches.size();
    if (plan.fused_of.size() != nl) plan.fused_of.assign(nl, -1);
    auto unit_of = [&](size_t i) {
        return plan.fused_of[i] < 0 ? uint32_t(i) : plan.fused[size_t(plan.fused_of[i])].leader;
    };
    std::vector<double> longest(nl, 0.0), total(nl, 0.0), alone(nl, 0.0);
    for (size_t i = 0; i < nl; ++i) {
        const Launch &L = plan.launches[i];
        const uint32_t u = unit_of(i);
        for (uint32_t t = L.task_begin; t < L.task_end; ++t) {
            const double c = task_cost(L.key, plan.tasks[t], plan.tuning);
            longest[u] = std::max(longest[u], c);
            total[u] += c;
        }
    }
    for (size_t i = 0; i < nl; ++i) alone[i] = std::max(longest[i], total[i] / kResidentGroups);
    plan.launch_order.resize(nl);
    std::iota(plan.launch_order.begin(), plan.launch_order.end(), 0u);
    std::stable_sort(plan.launch_order.begin(), plan.launch_order.end(), [&](uint32_t x, uint32_t y) {
        const bool mx = unit_of(x) != x, my = unit_of(y) != y;   // fused members other than the leader
        if (mx != my) return my;
        if (mx) return false;
        const Launch &a = plan.launches[x], &b = plan.launches[y];
        const bool wa = a.key.G == 64, wb = b.key.G == 64;
        if (wa != wb) return wa;
        if (wa) return longest[x] > longest[y];
        if (alone[x] != alone[y]) return alone[x] > alone[y];
        return a.padded_cells > b.padded_cells;
    });
    // With two or more streams the wavefront launches take streams of their own (round robin
    // over as many as there are wavefront launches, leaving at least one stream to striped
    // launches when there are any), so that a long wavefront workgroup never holds up the first
    // striped launches and wavefront launches of different variants run side by side; the
    // striped launches go round robin over the remaining streams first and then over all of them,
    // so that the later (smaller) ones queue behind the wavefront launches, which end well before
    // the large striped ones, instead of forming the pass's tail behind those (uneven 1M batch,
    // default tuning: 51 -> 37 ms per pass).
    size_t n_units = 0;
    for (size_t i = 0; i < nl; ++i) n_units += unit_of(i) == i;
    const uint32_t ns = std::max<uint32_t>(1, std::min<uint32_t>(plan.tuning.streams, uint32_t(std::max<size_t>(n_units, 1))));
    uint32_t n_wf = 0;
    for (const Launch &L : plan.launches) n_wf += L.key.G == 64;
    const uint32_t n_striped = uint32_t(n_units) - n_wf;
    const uint32_t wf_streams = ns == 1 || n_wf == 0 ? 0 : std::min(n_wf, n_striped ? ns - 1 : ns);
    plan.launch_stream.assign(nl, 0);
    uint32_t next_wf = 0, next_st = 0;
    for (size_t i = 0; i < nl; ++i) {
        const uint32_t li = plan.launch_order[i];
        if (unit_of(li) != li) {
            plan.launch_stream[li] = plan.launch_stream[unit_of(li)];   // placed after its leader
        } else if (plan.launches[li].key.G == 64 && wf_streams) {
            plan.launch_stream[li] = uint8_t(next_wf++ % wf_streams);
        } else {
            plan.launch_stream[li] = uint8_t((wf_streams + next_st++) % ns);
        }
    }
    // A fused launch queued behind another launch on its stream starts when that one ends and
    // takes the slots it leaves: it runs as many blocks as the GPU keeps resident instead of its
    // share of the pass (uneven 1M batch, launch form 4: the global R 4-12 launch behind the
    // local wavefront launch formed the pass's tail at one block per CU).
    std::vector<bool> stream_busy(ns, false);
    for (size_t i = 0; i < nl; ++i) {
        const uint32_t li = plan.launch_order[i];
        if (unit_of(li) != li) continue;
        const uint8_t s = plan.launch_stream[li];
        if (stream_busy[s] && plan.fused_of[li] >= 0) plan.fused[size_t(plan.fused_of[li])].share = 1.0;
        stream_busy[s] = true;
    }
}

int32_t max_substitution(const pa_config &cfg)
{
    if (cfg.score_kind == PA_SCORE_UNITARY) return std::max(cfg.match, cfg.mismatch);
    int32_t m = INT32_MIN;
    if (cfg.matrix)
        for (int32_t i = 0; i < cfg.matrix_dim * cfg.matrix_dim; ++i) m = std::max(m, cfg.matrix[i]);
    return m;
}

namespace {
int32_t min_substitution(const pa_config &cfg)
{
    if (cfg.score_kind == PA_SCORE_UNITARY) return std::min(cfg.match, cfg.mismatch);
    int32_t m = INT32_MAX;
    if (cfg.matrix)
        for (int32_t i = 0; i < cfg.matrix_dim * cfg.matrix_dim; ++i) m = std::min(m, cfg.matrix[i]);
    return m;
}
} // namespace

int32_t int16_local_bias(int32_t smin, int32_t go, int32_t ge)
{
    const int64_t g = std::min<int64_t>(go, 0), e = std::min<int64_t>(ge, 0);
    const int64_t low = std::min<int64_t>({int64_t(smin), -1, g + e});   // -1: the pad's score
    return int32_t(std::min<int64_t>(-low - e + 1, 32767));
}

int32_t int16_bias(const Plan &plan, bool local, uint32_t m, uint32_t n)
{
    if (local) return plan.lbias;
    const RangeBound rb = range_bound(plan, false, m, n);
    const int64_t low = rb.lower + std::min<int64_t>(plan.smin, -1);
    return int32_t(std::min<int64_t>(-low - std::min<int64_t>(plan.ge, 0) + 1, 32767));
}

int32_t int16_score_limit(const pa_config &cfg)
{
    const int64_t e = std::min<int64_t>(cfg.gap_extension, 0);
    return int32_t(kInt16Top - int16_local_bias(min_substitution(cfg), cfg.gap_open, cfg.gap_extension) + e * 81);
}

pa_status make_plan(const pa_config &cfg, const BatchView &b, const uint8_t *forced_width,
                    const std::vector<uint32_t> *only, Plan &plan, std::string &err, bool track_all,
                    bool export_dp, const float *pair_cost)
{
    if (cfg.method != PA_METHOD_GLOBAL && cfg.method != PA_METHOD_LOCAL) { err = "invalid method"; return PA_ERR_INVALID_ARGUMENT; }
    // The cell update drops V + go + ge from the vertical gap state (it never exceeds V + ge
    // when go <= 0), so both penalties must be non-positive.
    if (cfg.gap_extension > 0 || cfg.gap_open > 0) {
        err = "the device path needs gap_open <= 0 and gap_extension <= 0";
        return PA_ERR_UNSUPPORTED;
    }
    if (b.one_query && b.off1[1] < b.off1[0]) { err = "offsets must be non-decreasing"; return PA_ERR_INVALID_ARGUMENT; }
    for (uint64_t p = 0; p < b.n_pairs; ++p) {
        if ((!b.one_query && b.off1[p + 1] < b.off1[p]) || b.off2[p + 1] < b.off2[p]) {
            err = "offsets must be non-decreasing";
            return PA_ERR_INVALID_ARGUMENT;
        }
        if (b.pair_method && b.pair_method[p] > 1) { err = "invalid per-pair method"; return PA_ERR_INVALID_ARGUMENT; }
    }
    plan = Plan{};
    plan.tuning = resolve_tuning(cfg.tuning);
    plan.go = cfg.gap_open;
    plan.ge = cfg.gap_extension;
    int fc = cfg.first_column_free != 0, fr = cfg.first_row_free != 0;
    int lc = cfg.last_column_free != 0, lr = cfg.last_row_free != 0;
    // A free last row without a free last column is turned into a free last column by
    // exchanging the roles of seq1 and seq2 (the score is invariant under that exchange
    // together with first_row<->first_column and last_row<->last_column).
    bool swap_all = lr && !lc;
    if (swap_all) { std::swap(fc, fr); std::swap(lc, lr); }
    const bool symmetric = (fc == fr) && (lc == lr);
    plan.fc = fc; plan.fr = fr; plan.lc = lc; plan.lr = lr;

    Alphabet alpha;
    if (pa_status st = build_alphabet(cfg, b, plan, alpha, err); st != PA_OK) return st;
    plan.lbias = int16_local_bias(plan.smin, plan.go, plan.ge);
    if (plan.score == kScoreMatrix && plan.alphabet <= kPairTableMaxAlphabet) {
        const uint32_t K = plan.alphabet;
        for (int v = 0; v < kScoreShifts; ++v) {
            const std::vector<int16_t> &t = plan.table[v];
            std::vector<uint32_t> &pt = plan.ptable[v];
            pt.resize(size_t(K) * K * K);
            for (uint32_t a1 = 0; a1 < K; ++a1)
                for (uint32_t a2 = 0; a2 < K; ++a2)
                    for (uint32_t c = 0; c < K; ++c)
                        pt[(size_t(a1) * K + a2) * K + c] = (uint32_t(uint16_t(t[size_t(a2) * plan.stride + c])) << 16) |
                                                            uint32_t(uint16_t(t[size_t(a1) * plan.stride + c]));
        }
    }

    // Query profile: one query against many targets, with the query on the row side of every
    // pair (any orientation but the forced exchange for a free last row).  Built for the rows
    // the allowed shapes pad the query to, if it fits the LDS of a block.
    Tuning profile_tuning = plan.tuning;
    profile_tuning.rows_mask &= kProfileRowsMask;
    if (!profile_tuning.rows_mask) profile_tuning.rows_mask = kProfileRowsMask;
    const uint64_t qlen = b.one_query ? b.len1(0) : 0;
    if (b.one_query && !swap_all && qlen > 0 && qlen <= plan.tuning.wavefront_min_rows) {
        uint64_t rows = 0;
        for (uint32_t i = 0; i < 7; ++i)
            if ((profile_tuning.rows_mask >> i) & 1u) {
                const uint64_t stripe = 16 * (4 + 2 * i);
                rows = std::max(rows, (qlen + stripe - 1) / stripe * stripe);
            }
        const uint32_t K = plan.score == kScoreDna ? alpha.size + 1 : plan.alphabet;   // with the pad rank
        if (uint64_t(K) * rows * 2 <= kProfileMaxBytes) {
            plan.profile_rows = uint32_t(rows);
            plan.profile_ranks = K;
            const uint8_t *q = b.s1(0);
            for (int v = 0; v < kScoreShifts; ++v) {
                std::vector<uint32_t> &pf = plan.profile[v];
                pf.assign(size_t(K) * rows / 2, 0xFFFFFFFFu);   // -1 in both halves: pad rows, pad rank
                for (uint32_t c = 0; c + 1 < K; ++c)
                    for (uint64_t i = 0; i < qlen; ++i) {
                        const uint32_t a = uint32_t(alpha.code[q[i]]);
                        const int32_t sc = plan.score == kScoreDna ? (a == c ? cfg.match : cfg.mismatch)
                                                                   : int32_t(plan.table[0][a * plan.stride + c]);
                        const size_t e = size_t(c) * rows + i;
                        const uint32_t h = uint32_t(uint16_t(int16_t(sc - v * cfg.gap_extension)));
                        uint32_t &w = pf[e / 2];
                        w = (e & 1) ? (w & 0xFFFFu) | (h << 16) : (w & 0xFFFF0000u) | h;
                    }
            }
        }
    }
    const bool use_profile = !plan.profile[0].empty();

    const uint64_t N = b.n_pairs;
    plan.pairs.resize(N);
    std::vector<uint8_t> selected(N, only ? 0 : 1);
    if (only) for (uint32_t p : *only) selected[p] = 1;

    // Rows per lane of the striped kernel by DP rows m, tabulated once for the common lengths.
    auto rows_table = [&](const Tuning &t) {
        std::vector<uint8_t> tab(std::min<uint64_t>(uint64_t(t.wavefront_min_rows) + 1, 16384));
        parallel_for(tab.size(), [&](size_t m) {
            uint8_t G, R;
            choose_shape(uint32_t(m), t, false, G, R);
            tab[m] = R;
        });
        return tab;
    };
    const std::vector<uint8_t> rows_for_m = rows_table(plan.tuning);
    const std::vector<uint8_t> profile_rows_for = use_profile ? rows_table(profile_tuning) : std::vector<uint8_t>{};
    // Spread bound of the re-centred int16 wavefront tasks: one step (row or column) changes H
    // by at most smax + |go| + |ge| (+ the pad's 1 and |smin| for the gap and pad states).
    const int64_t recentre_step = std::max<int64_t>(plan.smax, 0) + std::max<int64_t>(-int64_t(plan.smin), 1) +
                                  std::abs(int64_t(plan.go)) + std::abs(int64_t(plan.ge));
    const bool recentre_ok = recentre_step * int64_t(kRecentreReach) < 16000;
    // Long pairs: the wavefront kernel spreads one pair over a workgroup of four waves, which
    // pays only where the pair's striped sweep would outlast the rest of the batch; where the
    // batch holds enough other work, a lane group of the striped kernel sweeps the pair faster.
    // A pair beyond wavefront_min_rows goes to the wavefront kernel when its modelled striped
    // cost exceeds wavefront_tail times the batch's striped work per resident lane group (a pair
    // expected in int16 lanes counts half, as int16 tasks hold two pairs; the orientation is taken
    // as rows = the shorter sequence), and always beyond kWavefrontAlwaysRows rows, where the
    // workgroup's four waves sweep two or more full rounds of stripes.
    double tail_cost = 0;
    if (!export_dp && plan.tuning.wavefront_tail > 0) {
        constexpr size_t kBlock = 1 << 14;
        const size_t nblk = (N + kBlock - 1) / kBlock;
        std::vector<double> part(nblk, 0.0);
        std::vector<uint8_t> has_long(nblk, 0);
        // the lane width the pair is expected to get below (re-centred long global pairs as int16)
        auto expect16 = [&](size_t p, bool local, uint32_t m, uint32_t n) {
            if (forced_width && forced_width[p]) return forced_width[p] == 16;
            if (cfg.lane_policy == PA_LANES_INT32 || -int64_t(plan.ge) * 81 > 16000) return false;
            if (!local && recentre_ok && m > plan.tuning.wavefront_min_rows) return true;
            const RangeBound rb = range_bound(plan, local, m, n);
            const bool lower16 = rb.lower > -32700;
            const bool upper16 = local ? int16_fits(rb, plan.lbias) : rb.upper < 32767;
            return cfg.lane_policy == PA_LANES_AUTO ? lower16 && upper16 : lower16;
        };
        parallel_for(nblk, [&](size_t k) {
            for (size_t p = k * kBlock; p < std::min<size_t>(N, (k + 1) * kBlock); ++p) {
                if (!selected[p]) continue;
                const bool local = b.pair_method ? b.pair_method[p] == PA_METHOD_LOCAL : cfg.method == PA_METHOD_LOCAL;
                const uint32_t l1 = uint32_t(b.len1(p)), l2 = uint32_t(b.off2[p + 1] - b.off2[p]);
                const uint32_t m = std::min(l1, l2), n = std::max(l1, l2);
                if (m == 0) continue;
                part[k] += striped_cost(m, n, plan.tuning) * (expect16(p, local, m, n) ? 0.5 : 1.0);
                has_long[k] |= m > plan.tuning.wavefront_min_rows;
            }
        }, 1);
        if (std::find(has_long.begin(), has_long.end(), 1) != has_long.end()) {
            const double total = std::accumulate(part.begin(), part.end(), 0.0);
            tail_cost = plan.tuning.wavefront_tail * total / kResidentGroups;
        }
    }
    // Per pair, in parallel: orientation, lane width and kernel variant.
    std::vector<VariantKey> pair_key(N);
    std::vector<uint8_t> speculative(N, 0);   // int16 without a proof against saturation
    std::atomic<bool> out_of_range{false};
    parallel_for(N, [&](size_t p) {
        const bool local = b.pair_method ? b.pair_method[p] == PA_METHOD_LOCAL : cfg.method == PA_METHOD_LOCAL;
        uint32_t l1 = uint32_t(b.len1(p)), l2 = uint32_t(b.off2[p + 1] - b.off2[p]);
        PairMeta &pm = plan.pairs[p];
        pm.local = local;
        // the query profile needs the query as the rows; the score does not depend on the
        // orientation when the end gaps are symmetric (always for the local method)
        bool swapped = use_profile ? false : local ? (l1 > l2) : (symmetric ? (l1 > l2) : swap_all);
        pm.swapped = swapped;
        pm.m = swapped ? l2 : l1;
        pm.n = swapped ? l1 : l2;
        pm.width = 0;
        if (!selected[p] || pm.m == 0 || pm.n == 0) return;
        const RangeBound rb = range_bound(plan, local, pm.m, pm.n);
        if (rb.lower < INT32_MIN / 2 || rb.upper > INT32_MAX / 2) { out_of_range = true; return; }
        bool lower16 = rb.lower > -32700;
        // every local int16 variant is a max3 variant (int16 values up to kInt16Top)
        bool upper16 = local ? int16_fits(rb, plan.lbias) : rb.upper < 32767;
        // long global pairs without per-row maxima run re-centred in the wavefront kernel: int16
        // whatever the absolute scores when the spread bound holds; otherwise int16 only with
        // the absolute bounds and the saturation tracking (which keeps absolute values)
        const bool wavefront = !export_dp && pm.m > plan.tuning.wavefront_min_rows &&
                               (tail_cost == 0 || pm.m > kWavefrontAlwaysRows ||
                                striped_cost(pm.m, pm.n, plan.tuning) > tail_cost);
        const bool rc_candidate = wavefront && !local && !(lc && lr);
        if (rc_candidate && recentre_ok) lower16 = upper16 = true;
        // The int16 kernels shift values by up to |ge| per row (17 rows) and per column step (64
        // steps); with a gap extension that large the shifted frames alone would leave int16,
        // so such configurations run in int32 under every policy.
        if (-int64_t(plan.ge) * 81 > 16000) lower16 = upper16 = false;
        uint8_t width;
        if (export_dp) width = 32;   // exported rows and columns are exact int32
        else if (forced_width && forced_width[p]) width = forced_width[p];
        else if (cfg.lane_policy == PA_LANES_INT32) width = 32;
        else if (cfg.lane_policy == PA_LANES_INT16 || cfg.lane_policy == PA_LANES_LEARNED) width = lower16 ? 16 : 32;
        else width = (lower16 && upper16) ? 16 : 32;
        if (width == 16 && !lower16) width = 32;   // the lower bound is never speculated on
        pm.width = width;
        speculative[p] = width == 16 && !upper16;
        VariantKey &key = pair_key[p];
        key.width = width;
        const bool profiled = use_profile && width == 16;
        if (export_dp) {
            key.G = 16;
            key.R = kExportRows;
        } else {
            const std::vector<uint8_t> &rows_for = profiled ? profile_rows_for : rows_for_m;
            if (pm.m < rows_for.size()) {
                key.G = 16;
                key.R = rows_for[pm.m];
            } else {
                choose_shape(pm.m, profiled ? profile_tuning : plan.tuning, wavefront, key.G, key.R);
            }
        }
        key.export_dp = export_dp ? 1 : 0;
        if (profiled) key.score = kScoreProfile;
        else key.score = (plan.score == kScoreMatrix && width == 16 && !plan.ptable[0].empty()) ? uint8_t(kScorePairs) : plan.score;
        key.local = local;
        key.track = (!local && ((width == 16 && !upper16) || track_all ||
                                (width == 16 && rc_candidate && !recentre_ok))) ? 1 : 0;
        key.rowmax = (!local && lc && lr) ? 1 : 0;
        // a global pair whose range does not fit the max3 variants' window keeps absolute int16
        // values with the running maximum (exact: its 16-bit bound is proven or tracked)
        if (!local && max3_variant(key) && !int16_fits(rb, int16_bias(plan, false, pm.m, pm.n))) key.track = 1;
    });
    if (out_of_range) { err = "scores may leave the int32 range"; return PA_ERR_RANGE; }
    // Group the pairs by variant: groups in VariantKey order, pairs in pair order within a group
    // (planning is deterministic).  Blocks of pairs are scanned in parallel: distinct keys, then
    // counts per group, then the items scattered to their places.
    constexpr size_t kPairBlock = 1 << 16;
    auto packed = [](const VariantKey &k) { uint64_t v; std::memcpy(&v, &k, sizeof v); return v; };
    const size_t nblk = (N + kPairBlock - 1) / kPairBlock;
    std::vector<std::vector<VariantKey>> blk_keys(nblk);
    parallel_for(nblk, [&](size_t k) {
        std::vector<uint64_t> seen_keys;
        for (size_t p = k * kPairBlock; p < std::min<size_t>(N, (k + 1) * kPairBlock); ++p) {
            if (plan.pairs[p].width == 0) continue;
            const uint64_t pk = packed(pair_key[p]);
            if (std::find(seen_keys.begin(), seen_keys.end(), pk) != seen_keys.end()) continue;
            seen_keys.push_back(pk);
            blk_keys[k].push_back(pair_key[p]);
        }
    }, 1);
    std::vector<VariantKey> keys;
    for (const auto &ks : blk_keys) keys.insert(keys.end(), ks.begin(), ks.end());
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const size_t ng = keys.size();
    std::vector<uint64_t> packed_keys(ng);
    for (size_t g = 0; g < ng; ++g) packed_keys[g] = packed(keys[g]);
    std::vector<uint32_t> gid(N);
    std::vector<uint64_t> blk_pos(nblk * ng, 0);   // per block and group: count, then start
    parallel_for(nblk, [&](size_t k) {
        uint32_t g = 0;
        for (size_t p = k * kPairBlock; p < std::min<size_t>(N, (k + 1) * kPairBlock); ++p) {
            if (plan.pairs[p].width == 0) continue;
            const uint64_t pk = packed(pair_key[p]);
            if (packed_keys[g] != pk)
                g = uint32_t(std::find(packed_keys.begin(), packed_keys.end(), pk) - packed_keys.begin());
            gid[p] = g;
            ++blk_pos[k * ng + g];
        }
    }, 1);
    std::vector<std::vector<Item>> items(ng);
    for (size_t g = 0; g < ng; ++g) {
        uint64_t total = 0;
        for (size_t k = 0; k < nblk; ++k) {
            const uint64_t c = blk_pos[k * ng + g];
            blk_pos[k * ng + g] = total;
            total += c;
        }
        items[g].resize(total);
    }
    struct Counts { uint64_t cells = 0, n16 = 0, n32 = 0, spec = 0, prof = 0, empty = 0; };
    std::vector<Counts> blk_counts(nblk);
    parallel_for(nblk, [&](size_t k) {
        Counts &c = blk_counts[k];
        for (size_t p = k * kPairBlock; p < std::min<size_t>(N, (k + 1) * kPairBlock); ++p) {
            const PairMeta &pm = plan.pairs[p];
            if (pm.width == 0) {
                if (selected[p]) ++c.empty;
                continue;
            }
            const uint32_t g = gid[p];
            const VariantKey &key = keys[g];
            const uint32_t rows = uint32_t(key.G) * key.R;
            items[g][blk_pos[k * ng + g]++] = Item{uint32_t(p), pm.m, pm.n, (pm.m + rows - 1) / rows};
            c.cells += uint64_t(pm.m) * pm.n;
            if (key.width == 16) ++c.n16; else ++c.n32;
            if (speculative[p]) ++c.spec;
            if (key.score == kScoreProfile) ++c.prof;
        }
    }, 1);
    for (const Counts &c : blk_counts) {
        plan.cells += c.cells;
        plan.n16 += c.n16;
        plan.n32 += c.n32;
        plan.n_speculative += c.spec;
        plan.n_profiled += c.prof;
        plan.n_empty += c.empty;
    }

    for (int c = 0; c < 256; ++c) plan.code[c] = alpha.code[c];
    plan.pad_rank = plan.alphabet ? plan.alphabet - 1 : 0;
    const uint64_t seq2_base = b.end1();   // the residue buffer holds seq1 bytes, then seq2 bytes
    // DP export output: dp_column values of every pair (seq1 side), then the dp_row values
    const uint64_t dp_total1 = b.one_query ? 0 : (b.off1[N] - b.off1[0]) + N;
    if (export_dp) {
        if (b.one_query) { err = "DP export needs one seq1 per pair"; return PA_ERR_INVALID_ARGUMENT; }
        plan.dp_entries = dp_total1 + (b.off2[N] - b.off2[0]) + N;
    }

    // Tasks of each group: int16 pairs packed two per task (similar stripe count and column
    // count side by side), longest tasks first so that the tail of a launch is made of short
    // tasks.
    using Pack = std::pair<int64_t, int64_t>;   // indices into the group's items; -1: no second half
    parallel_sort(items, [](const Item &a, const Item &c) {
        return std::tie(a.S, a.n, a.m, a.pair) > std::tie(c.S, c.n, c.m, c.pair);
    }, false);
    std::vector<std::vector<Pack>> group_packs(ng);
    parallel_for(ng, [&](size_t g) {
        const std::vector<Item> &it = items[g];
        std::vector<Pack> &packs = group_packs[g];
        if (keys[g].width == 16) {
            size_t i = 0;
            while (i < it.size()) {
                if (i + 1 < it.size() && (!keys[g].rowmax || it[i].n == it[i + 1].n)) {
                    packs.push_back({int64_t(i), int64_t(i + 1)});
                    i += 2;
                } else {
                    packs.push_back({int64_t(i), -1});
                    i += 1;
                }
            }
        } else {
            packs.resize(it.size());
            for (size_t i = 0; i < it.size(); ++i) packs[i] = {int64_t(i), -1};
        }
    }, 1);
    // work of a pack: stripes x (columns + G - 1), kept with the pack for the sort
    struct Weighed { uint64_t work; Pack pk; };
    std::vector<std::vector<Weighed>> weighed(ng);
    parallel_for(ng, [&](size_t g) {
        const std::vector<Item> &it = items[g];
        const uint32_t G = keys[g].G;
        weighed[g].resize(group_packs[g].size());
        for (size_t i = 0; i < group_packs[g].size(); ++i) {
            const Pack &pk = group_packs[g][i];
            uint32_t S = it[pk.first].S, n = it[pk.first].n;
            if (pk.second >= 0) { S = std::max(S, it[pk.second].S); n = std::max(n, it[pk.second].n); }
            weighed[g][i] = Weighed{uint64_t(S) * (n + G - 1), pk};
        }
    }, 1);
    parallel_sort(weighed, [](const Weighed &x, const Weighed &y) { return x.work > y.work; }, true);
    if (pair_cost) {
        // Predicted-cost order.  The lane groups of one wave sweep consecutive tasks in lockstep,
        // so a wave lasts as long as its longest task: the geometric order above keeps the tasks
        // of a wave alike, and the network's cost then orders whole waves (64 / G consecutive
        // tasks, keyed by the predicted cost of their costliest pair; one task per workgroup in
        // the wavefront kernel).  A last, partial wave stays last, so that the waves after it keep
        // their alignment.
        parallel_for(ng, [&](size_t g) {
            const std::vector<Item> &it = items[g];
            const size_t per_wave = keys[g].G >= 64 ? 1 : 64 / keys[g].G;
            std::vector<Weighed> &w = weighed[g];
            std::vector<std::pair<float, size_t>> waves;   // (key, first task)
            const size_t whole = w.size() / per_wave * per_wave;
            for (size_t lo = 0; lo < whole; lo += per_wave) {
                float c = 0.0f;
                for (size_t i = lo; i < std::min(w.size(), lo + per_wave); ++i) {
                    c = std::max(c, pair_cost[it[w[i].pk.first].pair]);
                    if (w[i].pk.second >= 0) c = std::max(c, pair_cost[it[w[i].pk.second].pair]);
                }
                waves.push_back({c, lo});
            }
            std::stable_sort(waves.begin(), waves.end(),
                             [](const auto &x, const auto &y) { return x.first > y.first; });
            std::vector<Weighed> out;
            out.reserve(w.size());
            for (const auto &wv : waves)
                for (size_t i = wv.second; i < wv.second + per_wave; ++i) out.push_back(w[i]);
            for (size_t i = whole; i < w.size(); ++i) out.push_back(w[i]);
            w.swap(out);
        }, 1);
    }
    parallel_for(ng, [&](size_t g) {
        for (size_t i = 0; i < weighed[g].size(); ++i) group_packs[g][i] = weighed[g][i].pk;
    }, 1);

    // Task records, in group order.  Chunks of packs are measured in parallel (code words,
    // boundary entries, lane-cells), their offsets are prefix sums in task order, and the
    // chunks are then filled in parallel.
    constexpr size_t kPackChunk = 1 << 13;
    struct Chunk { size_t g, lo, hi; uint64_t task, rowcodes, colcodes, bnd, padded; };
    std::vector<Chunk> chunks;
    for (size_t g = 0; g < ng; ++g)
        for (size_t lo = 0; lo < group_packs[g].size(); lo += kPackChunk)
            chunks.push_back(Chunk{g, lo, std::min(group_packs[g].size(), lo + kPackChunk), 0, 0, 0, 0, 0});
    auto task_shape = [&](size_t g, const Pack &pk, uint32_t &S, uint32_t &Nc) {
        const Item &a = items[g][pk.first];
        S = a.S;
        Nc = a.n;
        if (pk.second >= 0) { S = std::max(S, items[g][pk.second].S); Nc = std::max(Nc, items[g][pk.second].n); }
    };
    parallel_for(chunks.size(), [&](size_t c) {
        Chunk &ch = chunks[c];
        const VariantKey &key = keys[ch.g];
        const uint32_t G = key.G, rows = G * key.R;
        for (size_t i = ch.lo; i < ch.hi; ++i) {
            uint32_t S, Nc;
            task_shape(ch.g, group_packs[ch.g][i], S, Nc);
            const uint32_t cw = col_words(Nc, G);
            if (S > 1) ch.bnd += recentred(key) ? 2 * uint64_t(cw) : cw;
            if (key.score != kScoreProfile) ch.rowcodes += uint64_t(S) * rows / (key.score == kScorePairs ? 2 : 1);
            ch.colcodes += cw;
            ch.padded += uint64_t(S) * rows * Nc * (key.width == 16 ? 2 : 1);
        }
    }, 1);
    uint64_t n_rowcodes = 0, n_colcodes = 0, n_tasks = 0;
    for (Chunk &ch : chunks) {
        const uint64_t r = ch.rowcodes, cc = ch.colcodes, bn = ch.bnd;
        ch.task = n_tasks;
        ch.rowcodes = n_rowcodes;
        ch.colcodes = n_colcodes;
        ch.bnd = plan.bnd_entries;
        n_tasks += ch.hi - ch.lo;
        n_rowcodes += r;
        n_colcodes += cc;
        plan.bnd_entries += bn;
        plan.padded_cells += ch.padded;
    }
    plan.tasks.resize(n_tasks);
    plan.enc.resize(n_tasks);
    if (export_dp) plan.dp_base.resize(2 * n_tasks);
    parallel_for(chunks.size(), [&](size_t c) {
        const Chunk &ch = chunks[c];
        const VariantKey &key = keys[ch.g];
        const uint32_t G = key.G, rows = G * key.R;
        uint64_t rc = ch.rowcodes, cc = ch.colcodes, bn = ch.bnd, ti = ch.task;
        for (size_t i = ch.lo; i < ch.hi; ++i, ++ti) {
            const Pack &pk = group_packs[ch.g][i];
            const Item *h[2] = {&items[ch.g][pk.first], pk.second >= 0 ? &items[ch.g][pk.second] : nullptr};
            Task &t = plan.tasks[ti];
            t = Task{};
            task_shape(ch.g, pk, t.S, t.N);
            for (int q = 0; q < 2; ++q) {
                t.pair[q] = h[q] ? h[q]->pair : kNone;
                t.m[q] = h[q] ? h[q]->m : 0;
                t.n[q] = h[q] ? h[q]->n : 0;
                if (h[q] && !key.local && max3_variant(key))
                    t.bias |= uint32_t(int16_bias(plan, false, h[q]->m, h[q]->n)) << (16 * q);
            }
            t.row_off = uint32_t(std::min<uint64_t>(rc, 0xFFFFFFFFull));
            t.col_off = uint32_t(std::min<uint64_t>(cc, 0xFFFFFFFFull));
            t.bnd_off = kNone;
            const uint32_t cw = col_words(t.N, G);
            if (t.S > 1) {
                t.bnd_off = uint32_t(std::min<uint64_t>(bn, 0xFFFFFFFFull));
                bn += recentred(key) ? 2 * uint64_t(cw) : cw;
            }
            if (key.score != kScoreProfile) rc += uint64_t(t.S) * rows / (key.score == kScorePairs ? 2 : 1);
            cc += cw;
            EncodeTask &e = plan.enc[ti];
            e = EncodeTask{};
            for (int q = 0; q < 2; ++q) {
                if (!h[q]) continue;
                const uint32_t p = h[q]->pair;
                const uint64_t s1 = b.src1(p), s2 = seq2_base + b.off2[p];
                e.row_src[q] = plan.pairs[p].swapped ? s2 : s1;
                e.col_src[q] = plan.pairs[p].swapped ? s1 : s2;
            }
            if (export_dp) {
                // our rows are seq1 unless the pair is exchanged (then seq2)
                const uint32_t p = h[0]->pair;
                const uint64_t c1 = b.off1[p] - b.off1[0] + p, c2 = dp_total1 + (b.off2[p] - b.off2[0]) + p;
                const bool sw = plan.pairs[p].swapped;
                plan.dp_base[2 * ti] = sw ? c2 : c1;       // H(0, n): the task's last column
                plan.dp_base[2 * ti + 1] = sw ? c1 : c2;   // H(m, 0): the task's last row
            }
        }
    }, 1);
    for (size_t g = 0; g < ng; ++g) {
        Launch L{keys[g], 0, 0, 0};
        bool first = true;
        for (const Chunk &ch : chunks) {
            if (ch.g != g) continue;
            if (first) L.task_begin = uint32_t(ch.task);
            first = false;
            L.task_end = uint32_t(ch.task + ch.hi - ch.lo);
            L.padded_cells += ch.padded;
        }
        if (L.task_end > L.task_begin) plan.launches.push_back(L);
    }
    if (n_rowcodes > 0xFFFFFFF0ull || n_colcodes > 0xFFFFFFF0ull || plan.bnd_entries > 0x7FFFFFF0ull) {
        err = "batch too large for one plan; split it";
        return PA_ERR_INVALID_ARGUMENT;
    }

    plan.n_rowcodes = n_rowcodes;
    plan.n_colcodes = n_colcodes;
    fuse_launches(plan, pair_cost);
    order_launches(plan);
    return PA_OK;
}

} // namespace pa
