// Host planner checks (no GPU): make_plan on seeded random batches, against what a plan must
// satisfy whatever its packing choices:
//   * every pair with two non-empty sequences sits in exactly one task half, empty pairs in none;
//   * each half fits its task (m <= S*G*R, n <= N), int32 tasks have one half, keys agree with the
//     pair's width and method, launches partition the tasks into runs of one key;
//   * code-word, boundary and DP-export offsets are the running sums the kernels assume;
//   * the encoder's residue sources point at the pair's sequences in the uploaded buffer;
//   * cells and the int16/int32 counts add up; the residue codes equal a byte-by-byte scan
//     (the planner's vectorised alphabet scan, bytes >= 0x80 and ragged block ends included);
//   * planning twice gives the same plan.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "pa_plan.hpp"

using namespace pa;

static int failures = 0;
#define CHECK(cond, ...)                                          \
    do {                                                          \
        if (!(cond)) {                                            \
            if (failures < 20) {                                  \
                std::printf("FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
                std::printf(__VA_ARGS__);                         \
                std::printf("\n");                                \
            }                                                     \
            ++failures;                                           \
        }                                                         \
    } while (0)

struct Batch {
    std::vector<uint8_t> s1, s2, method;
    std::vector<uint64_t> o1{0}, o2{0};
    bool one_query = false;
    BatchView view() const
    {
        return BatchView{s1.data(), o1.data(), s2.data(), o2.data(), o2.size() - 1, method.empty() ? nullptr : method.data(),
                         one_query};
    }
};

static Batch random_batch(std::mt19937_64 &rng, int n, const char *alphabet, int max_len, bool one_query, bool long_pairs,
                          bool odd_bytes)
{
    Batch b;
    b.one_query = one_query;
    const size_t k = std::strlen(alphabet);
    auto seq = [&](std::vector<uint8_t> &out, int len) {
        for (int i = 0; i < len; ++i) out.push_back(uint8_t(alphabet[rng() % k]));
        if (odd_bytes && len > 3 && rng() % 50 == 0) out[out.size() - 1 - rng() % 3] = uint8_t(0x80 + rng() % 8);
    };
    if (one_query) {
        seq(b.s1, 1 + int(rng() % max_len));
        b.o1.push_back(b.s1.size());
    }
    for (int p = 0; p < n; ++p) {
        int l1 = int(rng() % (max_len + 1)), l2 = int(rng() % (max_len + 1));
        if (p % 37 == 0) l1 = 0;
        if (p % 41 == 0) l2 = 0;
        if (long_pairs && p % 97 == 5) l1 = 3100 + int(rng() % 2000);
        if (!one_query) {
            seq(b.s1, l1);
            b.o1.push_back(b.s1.size());
        }
        seq(b.s2, l2);
        b.o2.push_back(b.s2.size());
        b.method.push_back(uint8_t(rng() % 2));
    }
    if (b.s1.empty()) b.s1.push_back('A');
    if (b.s2.empty()) b.s2.push_back('A');
    return b;
}

static uint64_t fnv(const void *p, size_t n, uint64_t h)
{
    const uint8_t *c = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

static uint64_t digest(const Plan &p)
{
    uint64_t h = 1469598103934665603ull;
    h = fnv(p.pairs.data(), p.pairs.size() * sizeof(PairMeta), h);
    h = fnv(p.tasks.data(), p.tasks.size() * sizeof(Task), h);
    h = fnv(p.enc.data(), p.enc.size() * sizeof(EncodeTask), h);
    h = fnv(p.dp_base.data(), p.dp_base.size() * 8, h);
    for (const Launch &L : p.launches) h = fnv(&L, sizeof L, h);
    h = fnv(p.launch_order.data(), p.launch_order.size() * 4, h);
    return fnv(p.launch_stream.data(), p.launch_stream.size(), h);
}

static void check_plan(const char *label, const pa_config &cfg, const Batch &b, const uint8_t *forced = nullptr,
                       const std::vector<uint32_t> *only = nullptr, bool export_dp = false)
{
    const BatchView v = b.view();
    Plan p;
    std::string err;
    const pa_status st = make_plan(cfg, v, forced, only, p, err, false, export_dp);
    if (st == PA_ERR_INVALID_SYMBOL) {   // a byte outside the matrix alphabet: expected for odd bytes
        std::printf("  %s: rejected (%s)\n", label, err.c_str());
        return;
    }
    CHECK(st == PA_OK, "%s: status %d %s", label, st, err.c_str());
    if (st != PA_OK) return;
    const uint64_t N = v.n_pairs;
    std::vector<bool> selected(N, only == nullptr);
    if (only)
        for (uint32_t q : *only) selected[q] = true;

    // residue codes against a byte-by-byte scan
    bool seen[256] = {};
    for (uint64_t i = v.off1[0]; i < v.end1(); ++i) seen[b.s1[i]] = true;
    for (uint64_t i = 0; i < v.off2[N]; ++i) seen[b.s2[i]] = true;
    if (cfg.score_kind == PA_SCORE_UNITARY) {
        int k = 0;
        for (int c = 0; c < 256; ++c) CHECK(p.code[c] == (seen[c] ? k++ : -1), "%s: code of byte %d", label, c);
    }

    // every non-empty selected pair in exactly one half
    std::vector<int> count(N, 0);
    uint64_t rowcodes = 0, colcodes = 0, bnd = 0, cells = 0, n16 = 0, n32 = 0;
    for (const Launch &L : p.launches) {
        const VariantKey &key = L.key;
        const uint32_t rows = uint32_t(key.G) * key.R;
        CHECK(L.task_begin < L.task_end && L.task_end <= p.tasks.size(), "%s: launch range", label);
        for (uint32_t t = L.task_begin; t < L.task_end; ++t) {
            const Task &tk = p.tasks[t];
            CHECK(tk.row_off == uint32_t(rowcodes) && tk.col_off == uint32_t(colcodes), "%s: code offsets of task %u", label, t);
            const uint32_t cw = col_words(tk.N, key.G);
            if (tk.S > 1) {
                CHECK(tk.bnd_off == uint32_t(bnd), "%s: boundary offset of task %u", label, t);
                bnd += recentred(key) ? 2 * uint64_t(cw) : cw;   // exact int32 rows when re-centred
            } else {
                CHECK(tk.bnd_off == kNone, "%s: single-stripe task %u has a boundary", label, t);
            }
            if (key.score != kScoreProfile) rowcodes += uint64_t(tk.S) * rows / (key.score == kScorePairs ? 2 : 1);
            colcodes += cw;
            CHECK(key.width == 16 || tk.pair[1] == kNone, "%s: int32 task %u with two halves", label, t);
            for (int q = 0; q < 2; ++q) {
                if (tk.pair[q] == kNone) continue;
                const uint32_t pr = tk.pair[q];
                CHECK(pr < N, "%s: pair index", label);
                if (pr >= N) continue;
                ++count[pr];
                const PairMeta &pm = p.pairs[pr];
                CHECK(pm.width == key.width, "%s: width of pair %u", label, pr);
                CHECK(pm.local == key.local, "%s: method of pair %u", label, pr);
                CHECK(tk.m[q] == pm.m && tk.n[q] == pm.n && pm.m <= uint64_t(tk.S) * rows && pm.n <= tk.N,
                      "%s: shape of pair %u", label, pr);
                const uint64_t l1 = v.len1(pr), l2 = v.off2[pr + 1] - v.off2[pr];
                CHECK((pm.swapped ? l2 : l1) == pm.m && (pm.swapped ? l1 : l2) == pm.n, "%s: orientation of pair %u", label, pr);
                const uint64_t s1 = v.src1(pr), s2 = v.end1() + v.off2[pr];
                const EncodeTask &e = p.enc[t];
                CHECK(e.row_src[q] == (pm.swapped ? s2 : s1) && e.col_src[q] == (pm.swapped ? s1 : s2),
                      "%s: residue sources of pair %u", label, pr);
                if (export_dp && q == 0) {
                    const uint64_t c1 = v.off1[pr] - v.off1[0] + pr;
                    const uint64_t c2 = (v.off1[N] - v.off1[0]) + N + (v.off2[pr] - v.off2[0]) + pr;
                    CHECK(p.dp_base[2 * t] == (pm.swapped ? c2 : c1) && p.dp_base[2 * t + 1] == (pm.swapped ? c1 : c2),
                          "%s: export offsets of pair %u", label, pr);
                }
                // int16 three-way-maximum variants: every operand of a three-way maximum biased into
                // [0, kInt16Top] -- the lower bound less one substitution score lifted to >= 0, the
                // upper bound (drift included) at most kInt16Top; global halves carry their bias
                const uint32_t hb = (tk.bias >> (16 * q)) & 0xFFFFu;
                if (max3_variant(key)) {
                    const RangeBound rb = range_bound(p, pm.local, pm.m, pm.n);
                    const int64_t b = pm.local ? p.lbias : int64_t(hb);
                    CHECK(pm.local ? hb == 0 : hb == uint32_t(int16_bias(p, false, pm.m, pm.n)),
                          "%s: bias of pair %u", label, pr);
                    // local: H >= 0, so an operand is >= min(smin, pad -1, go + ge); global: >= the
                    // lower bound of H and the gap states, or one substitution score below it
                    const int64_t low = pm.local ? std::min<int64_t>({p.smin, -1, int64_t(p.go) + p.ge})
                                                 : rb.lower + std::min(p.smin, -1);
                    CHECK(low + b >= 0, "%s: biased lower bound of pair %u", label, pr);
                    if (!pm.local) CHECK(rb.upper + b <= kInt16Top, "%s: biased upper bound of pair %u", label, pr);
                } else {
                    CHECK(hb == 0, "%s: bias of a pair outside the max3 variants", label);
                }
                cells += uint64_t(pm.m) * pm.n;
                if (key.width == 16) ++n16; else ++n32;
            }
        }
    }
    CHECK(rowcodes == p.n_rowcodes && colcodes == p.n_colcodes && bnd == p.bnd_entries, "%s: code totals", label);
    CHECK(cells == p.cells && n16 == p.n16 && n32 == p.n32, "%s: cell and width counts", label);
    for (uint64_t q = 0; q < N; ++q) {
        const bool nonempty = v.len1(q) > 0 && v.off2[q + 1] > v.off2[q];
        CHECK(count[q] == ((selected[q] && nonempty) ? 1 : 0), "%s: pair %llu in %d halves", label,
              (unsigned long long)q, count[q]);
        if (forced && selected[q] && nonempty && forced[q]) CHECK(p.pairs[q].width == forced[q], "%s: forced width", label);
    }
    std::set<uint32_t> order(p.launch_order.begin(), p.launch_order.end());
    CHECK(order.size() == p.launches.size() && p.launch_stream.size() == p.launches.size(), "%s: launch order", label);

    Plan again;
    make_plan(cfg, v, forced, only, again, err, false, export_dp);
    CHECK(digest(again) == digest(p), "%s: planning is not deterministic", label);
    std::printf("  %s: %llu pairs, %zu tasks, %zu launches\n", label, (unsigned long long)N, p.tasks.size(), p.launches.size());
}

// Long global pairs: re-centred int16 exactly when the spread bound holds (no saturation
// tracking), otherwise int32 or int16 on the absolute bounds, then with the saturation tracking.
static void check_recentring()
{
    std::mt19937_64 rng(7);
    Batch b;
    for (int p = 0; p < 40; ++p) {
        const int l1 = 3200 + int(rng() % 4000), l2 = 3200 + int(rng() % 4000);
        for (int i = 0; i < l1; ++i) b.s1.push_back("ACGT"[rng() % 4]);
        for (int i = 0; i < l2; ++i) b.s2.push_back("ACGT"[rng() % 4]);
        b.o1.push_back(b.s1.size());
        b.o2.push_back(b.s2.size());
    }
    struct Case { int match, mismatch, go, ge, policy; bool want_rc; uint8_t want_width; };
    const Case cases[] = {
        {1, -4, -6, -1, PA_LANES_AUTO, true, 16},     // C = 12: 14,592 < 16,000
        {1, -4, -6, -1, PA_LANES_INT32, false, 32},
        {4, -5, -10, -1, PA_LANES_AUTO, false, 16},   // C = 20: beyond the bound; absolute bounds hold
        {4, -5, -10, -1, PA_LANES_INT16, false, 16},  // tracked int16 (absolute values)
    };
    for (const Case &c : cases) {
        pa_config cfg{};
        cfg.match = c.match;
        cfg.mismatch = c.mismatch;
        cfg.gap_open = c.go;
        cfg.gap_extension = c.ge;
        cfg.lane_policy = c.policy;
        Plan p;
        std::string err;
        const BatchView v = b.view();
        CHECK(make_plan(cfg, v, nullptr, nullptr, p, err) == PA_OK, "recentring plan: %s", err.c_str());
        for (const Launch &L : p.launches) {
            CHECK(L.key.G == 64, "long pairs in the wavefront kernel");
            CHECK(L.key.width == c.want_width, "width %d, expected %d (policy %d)", L.key.width, c.want_width, c.policy);
            CHECK(recentred(L.key) == c.want_rc, "re-centred %d, expected %d", int(recentred(L.key)), int(c.want_rc));
            if (L.key.width == 16 && !c.want_rc) CHECK(L.key.track == 1, "non-re-centred int16 keeps the tracking");
        }
    }
    std::printf("  re-centring choices: checked\n");
}

// A gap extension so large that the int16 kernels' drift frames (|ge| per row and per column
// step) alone would leave int16: every pair runs in int32, under every lane policy.
static void check_large_gap_extension()
{
    std::mt19937_64 rng(11);
    Batch b;
    for (int p = 0; p < 30; ++p) {
        const int l1 = 20 + int(rng() % (p == 0 ? 4000 : 300)), l2 = 20 + int(rng() % 300);
        for (int i = 0; i < l1; ++i) b.s1.push_back("ACGT"[rng() % 4]);
        for (int i = 0; i < l2; ++i) b.s2.push_back("ACGT"[rng() % 4]);
        b.o1.push_back(b.s1.size());
        b.o2.push_back(b.s2.size());
    }
    for (int ge : {-600, -3000}) {
        for (int method : {PA_METHOD_GLOBAL, PA_METHOD_LOCAL}) {
            for (int policy : {PA_LANES_AUTO, PA_LANES_INT16}) {
                pa_config cfg{};
                cfg.method = method;
                cfg.match = 4;
                cfg.mismatch = -5;
                cfg.gap_open = -10;
                cfg.gap_extension = ge;
                cfg.lane_policy = policy;
                Plan p;
                std::string err;
                const BatchView v = b.view();
                CHECK(make_plan(cfg, v, nullptr, nullptr, p, err) == PA_OK, "large ge plan: %s", err.c_str());
                for (const Launch &L : p.launches)
                    CHECK(L.key.width == 32, "ge %d method %d policy %d: width %d", ge, method, policy, L.key.width);
            }
        }
    }
    std::printf("  large gap extension: int32 only\n");
}

// Long pairs go to the wavefront kernel only where their striped sweep would be the batch's
// tail: exactly the pairs beyond wavefront_min_rows whose modelled striped cost exceeds
// wavefront_tail x (the batch's striped work, pairs expected in int16 lanes counting half) /
// resident lane groups (every pair here is int16).  Alone, every long pair is such a tail; with a
// tail multiple that puts the threshold at the median long pair, half of them are; a negative
// multiple sends them all.
static void check_wavefront_tail()
{
    std::mt19937_64 rng(13);
    Batch b;
    for (int p = 0; p < 60; ++p) {
        const int l1 = p % 3 == 0 ? 200 + int(rng() % 800) : 3100 + int(rng() % 4000);
        const int l2 = l1 + int(rng() % 200) - 100;
        for (int i = 0; i < l1; ++i) b.s1.push_back("ACGT"[rng() % 4]);
        for (int i = 0; i < l2; ++i) b.s2.push_back("ACGT"[rng() % 4]);
        b.o1.push_back(b.s1.size());
        b.o2.push_back(b.s2.size());
    }
    pa_config cfg{};
    cfg.method = PA_METHOD_LOCAL;
    cfg.match = 2;
    cfg.mismatch = -3;
    cfg.gap_open = -5;
    cfg.gap_extension = -2;
    const BatchView v = b.view();
    const Tuning defaults;
    std::vector<double> cost(60), long_cost;
    double total = 0;
    for (int p = 0; p < 60; ++p) {
        const uint32_t l1 = uint32_t(b.o1[p + 1] - b.o1[p]), l2 = uint32_t(b.o2[p + 1] - b.o2[p]);
        cost[p] = striped_cost(std::min(l1, l2), std::max(l1, l2), defaults);
        total += cost[p];
        if (std::min(l1, l2) > defaults.wavefront_min_rows) long_cost.push_back(cost[p]);
    }
    std::sort(long_cost.begin(), long_cost.end());
    const double median = long_cost[long_cost.size() / 2];
    // wavefront_tail 0 (the default), a multiple placing the threshold just below the median long
    // pair, and -1 (every long pair)
    const float tails[] = {0.0f, float(median * 0.999 * kResidentGroups / (total * 0.5)), -1.0f};
    for (float tail : tails) {
        pa_tuning t{};
        t.wavefront_tail = tail;
        cfg.tuning = &t;
        Plan p;
        std::string err;
        CHECK(make_plan(cfg, v, nullptr, nullptr, p, err) == PA_OK, "tail plan: %s", err.c_str());
        const double threshold = tail > 0 ? tail * total * 0.5 / kResidentGroups : 0.0;
        size_t n_wf = 0;
        for (const Launch &L : p.launches)
            for (uint32_t k = L.task_begin; k < L.task_end; ++k)
                for (int h = 0; h < 2; ++h) {
                    const uint32_t q = p.tasks[k].pair[h];
                    if (q == kNone) continue;
                    const bool want = p.pairs[q].m > defaults.wavefront_min_rows &&
                                      (p.pairs[q].m > kWavefrontAlwaysRows || cost[q] > threshold);
                    CHECK((L.key.G == 64) == want, "tail %g pair %u (m %u): G %d", tail, q, p.pairs[q].m, L.key.G);
                    n_wf += L.key.G == 64;
                }
        if (tail > 0) CHECK(n_wf == long_cost.size() - long_cost.size() / 2, "tail %g: %zu wavefront pairs", tail, n_wf);
        else CHECK(n_wf == long_cost.size(), "tail %g: %zu wavefront pairs of %zu", tail, n_wf, long_cost.size());
    }
    // beyond kWavefrontAlwaysRows rows a pair goes to the wavefront kernel whatever the tail rule
    Batch big;
    for (int i = 0; i < int(kWavefrontAlwaysRows) + 50; ++i) big.s1.push_back("ACGT"[rng() % 4]);
    big.s2 = big.s1;
    big.o1.push_back(big.s1.size());
    big.o2.push_back(big.s2.size());
    pa_tuning never{};
    never.wavefront_tail = 1e9f;
    cfg.tuning = &never;
    Plan pb;
    std::string err;
    CHECK(make_plan(cfg, big.view(), nullptr, nullptr, pb, err) == PA_OK, "always plan: %s", err.c_str());
    CHECK(pb.launches.size() == 1 && pb.launches[0].key.G == 64, "a pair of %u rows in the wavefront kernel",
          kWavefrontAlwaysRows + 50);
    std::printf("  wavefront tail rule: checked\n");
}

// Predicted-cost task order (pa_tuning.cost_order): the lane groups of a wave sweep 64 / G
// consecutive tasks in lockstep, so the order may move whole waves only.  Every aligned run of
// 64 / G tasks of a striped launch holds the same pairs as an aligned run of the geometric order,
// and the runs come in non-increasing order of their costliest pair's predicted cost.
static void check_cost_order()
{
    std::mt19937_64 rng(29);
    Batch b = random_batch(rng, 30000, "ACGT", 600, false, true, false);
    pa_config cfg{};
    cfg.match = 2;
    cfg.mismatch = -3;
    cfg.gap_open = -5;
    cfg.gap_extension = -2;
    const BatchView v = b.view();
    const size_t n = b.o2.size() - 1;
    std::vector<float> cost(n);
    for (float &c : cost) c = float(rng() % 100000);
    Plan geo, by_cost;
    std::string err;
    CHECK(make_plan(cfg, v, nullptr, nullptr, geo, err) == PA_OK, "geometric plan: %s", err.c_str());
    CHECK(make_plan(cfg, v, nullptr, nullptr, by_cost, err, false, false, cost.data()) == PA_OK, "cost plan: %s",
          err.c_str());
    CHECK(geo.launches.size() == by_cost.launches.size(), "launch counts %zu %zu", geo.launches.size(),
          by_cost.launches.size());
    size_t runs = 0;
    for (size_t l = 0; l < std::min(geo.launches.size(), by_cost.launches.size()); ++l) {
        const Launch &A = geo.launches[l], &B = by_cost.launches[l];
        CHECK(A.task_begin == B.task_begin && A.task_end == B.task_end, "launch %zu ranges", l);
        const uint32_t per = A.key.G >= 64 ? 1 : 64 / A.key.G;
        auto run_pairs = [&](const Plan &p, uint32_t k0) {
            std::vector<uint32_t> s;
            for (uint32_t k = k0; k < std::min(k0 + per, A.task_end); ++k)
                for (int h = 0; h < 2; ++h) s.push_back(p.tasks[k].pair[h]);
            std::sort(s.begin(), s.end());
            return s;
        };
        auto run_key = [&](uint32_t k0) {
            float c = 0.0f;
            for (uint32_t k = k0; k < std::min(k0 + per, A.task_end); ++k)
                for (int h = 0; h < 2; ++h)
                    if (by_cost.tasks[k].pair[h] != kNone) c = std::max(c, cost[by_cost.tasks[k].pair[h]]);
            return c;
        };
        std::set<std::vector<uint32_t>> geo_runs;
        for (uint32_t k0 = A.task_begin; k0 < A.task_end; k0 += per) geo_runs.insert(run_pairs(geo, k0));
        float prev = INFINITY;
        for (uint32_t k0 = B.task_begin; k0 < B.task_end; k0 += per, ++runs) {
            CHECK(geo_runs.count(run_pairs(by_cost, k0)) == 1, "launch %zu: run at task %u is not a geometric run", l,
                  k0);
            if (k0 + per > B.task_end) break;   // a partial last run stays last
            const float c = run_key(k0);
            CHECK(c <= prev, "launch %zu: run at task %u costs %g after %g", l, k0, double(c), double(prev));
            prev = c;
        }
    }
    // the wave items of fused launches follow the predicted cost too
    size_t items = 0;
    for (const FusedLaunch &F : by_cost.fused) {
        float prev = INFINITY;
        for (uint32_t i = F.items_begin; i < F.items_end; ++i, ++items) {
            const uint32_t it = by_cost.fused_items[i];
            const uint32_t n = ((it >> 27) & 3u) + 1, first = it & (kFusedMaxTasks - 1);
            float c = 0.0f;
            for (uint32_t k = first; k < first + n; ++k)
                for (int h = 0; h < 2; ++h)
                    if (by_cost.tasks[k].pair[h] != kNone) c = std::max(c, cost[by_cost.tasks[k].pair[h]]);
            CHECK(c <= prev, "fused item %u: predicted cost %g after %g", i, double(c), double(prev));
            prev = c;
        }
    }
    CHECK(items > 0, "no fused wave items in the cost-ordered plan");
    std::printf("  cost order: %zu runs of whole waves and %zu fused wave items checked\n", runs, items);
}

// Fused launches (pa_tuning.persistent = 2): the int16 local striped launches that differ only in R
// form one fused launch; its wave items cover every task of every member exactly once, in runs of
// at most four consecutive tasks of one member (its R in the item), in non-increasing order of
// the modelled time of their costliest task; leaders come before the other members in the launch
// order, which share their leader's stream; global and int32 launches stay unfused.
static void check_fused_launches()
{
    std::mt19937_64 rng(31);
    Batch b = random_batch(rng, 40000, "ACGT", 600, false, true, false);   // per-pair methods
    pa_config cfg{};
    cfg.match = 2;
    cfg.mismatch = -3;
    cfg.gap_open = -5;
    cfg.gap_extension = -2;
    cfg.lane_policy = PA_LANES_INT16;
    pa_tuning t{};
    t.persistent = 2;
    cfg.tuning = &t;
    Plan p;
    std::string err;
    CHECK(make_plan(cfg, b.view(), nullptr, nullptr, p, err) == PA_OK, "fused plan: %s", err.c_str());
    const size_t nl = p.launches.size();
    CHECK(p.fused_of.size() == nl, "fused_of size %zu of %zu", p.fused_of.size(), nl);
    CHECK(!p.fused.empty(), "no fused launch among %zu launches", nl);
    for (size_t l = 0; l < nl && l < p.fused_of.size(); ++l) {
        const VariantKey &k = p.launches[l].key;
        if (p.fused_of[l] < 0) {
            // an unfused launchable launch of the local method (form 2 fuses no global launch)
            // has no other launch of its group
            if (!fused_launchable(k) || !k.local) continue;
            for (size_t o = 0; o < nl; ++o) {
                VariantKey a = k, c = p.launches[o].key;
                a.R = c.R = 0;
                CHECK(o == l || !(a == c) || p.fused_of[o] >= 0, "launches %zu and %zu differ only in R, unfused", l, o);
            }
            continue;
        }
        CHECK(fused_launchable(k) && k.local, "launch %zu (width %d, local %d) fused", l, k.width, k.local);
        const FusedLaunch &F = p.fused[size_t(p.fused_of[l])];
        VariantKey a = k, c = p.launches[F.leader].key;
        a.R = c.R = 0;
        CHECK(a == c, "launch %zu fused with a leader of another variant", l);
    }
    for (size_t f = 0; f < p.fused.size(); ++f) {
        const FusedLaunch &F = p.fused[f];
        std::vector<uint8_t> seen(p.tasks.size(), 0);
        double prev = INFINITY;
        for (uint32_t i = F.items_begin; i < F.items_end; ++i) {
            const uint32_t it = p.fused_items[i];
            const uint32_t R = 4 + 2 * (it >> 29), n = ((it >> 27) & 3u) + 1, first = it & (kFusedMaxTasks - 1);
            size_t owner = nl;
            for (size_t l = 0; l < nl; ++l)
                if (p.fused_of[l] == int32_t(f) && first >= p.launches[l].task_begin && first < p.launches[l].task_end)
                    owner = l;
            CHECK(owner < nl, "item %u: task %u in no member launch", i, first);
            if (owner == nl) continue;
            const Launch &L = p.launches[owner];
            CHECK(L.key.R == R, "item %u: R %u in a launch of R %d", i, R, L.key.R);
            CHECK(first + n <= L.task_end && (first - L.task_begin) % 4 == 0, "item %u: tasks %u+%u", i, first, n);
            double cost = 0;
            for (uint32_t k = first; k < std::min(first + n, L.task_end); ++k) {
                CHECK(seen[k] == 0, "task %u in two items", k);
                seen[k] = 1;
                cost = std::max(cost, task_cost(L.key, p.tasks[k], p.tuning));
            }
            CHECK(cost <= prev, "item %u: modelled time %g after %g", i, cost, prev);
            prev = cost;
        }
        for (size_t l = 0; l < nl; ++l)
            if (p.fused_of[l] == int32_t(f))
                for (uint32_t k = p.launches[l].task_begin; k < p.launches[l].task_end; ++k)
                    CHECK(seen[k] == 1, "fused launch %zu: task %u of launch %zu in no item", f, k, l);
    }
    std::vector<size_t> pos(nl);
    for (size_t i = 0; i < nl; ++i) pos[p.launch_order[i]] = i;
    for (size_t l = 0; l < nl; ++l) {
        if (p.fused_of[l] < 0) continue;
        const uint32_t leader = p.fused[size_t(p.fused_of[l])].leader;
        CHECK(l == leader || pos[leader] < pos[l], "member %zu ordered before its leader %u", l, leader);
        CHECK(p.launch_stream[l] == p.launch_stream[leader], "member %zu on another stream than its leader", l);
    }
    // the default form fuses the local launches of both score kinds; forms 1 and 3 fuse nothing
    static const char *kAmino = "ARNDCQEGHILKMFPSTWYVX";
    std::vector<int32_t> matrix(21 * 21);
    for (int i = 0; i < 21; ++i)
        for (int j = 0; j < 21; ++j) matrix[i * 21 + j] = i == j ? 5 : -1 - (i + j) % 3;
    pa_config prot{};
    prot.method = PA_METHOD_LOCAL;
    prot.gap_open = -11;
    prot.gap_extension = -1;
    prot.score_kind = PA_SCORE_MATRIX;
    prot.matrix_dim = 21;
    prot.matrix = matrix.data();
    prot.matrix_symbols = reinterpret_cast<const uint8_t *>(kAmino);
    Batch pb = random_batch(rng, 20000, kAmino, 500, false, false, false);
    for (uint32_t form : {0u, 1u, 3u}) {
        for (pa_config *c : {&prot, &cfg}) {
            pa_tuning tf{};
            tf.persistent = form;
            c->tuning = &tf;
            Plan q;
            const Batch &bb = c == &prot ? pb : b;
            CHECK(make_plan(*c, bb.view(), nullptr, nullptr, q, err) == PA_OK, "form %u plan: %s", form, err.c_str());
            const bool want = form == 0;
            CHECK(q.fused.empty() != want, "form %u, %s: %zu fused launches", form, c == &prot ? "protein" : "DNA",
                  q.fused.size());
            for (const FusedLaunch &F : q.fused)
                CHECK(q.launches[F.leader].key.score == (c == &prot ? kScorePairs : kScoreDna), "form %u", form);
        }
    }
    // form 4 also fuses the global method's launches, in the two ranges of rows per lane
    {
        pa_tuning t4{};
        t4.persistent = 4;
        cfg.tuning = &t4;
        Plan q;
        CHECK(make_plan(cfg, b.view(), nullptr, nullptr, q, err) == PA_OK, "form 4 plan: %s", err.c_str());
        size_t global_groups = 0, global_members = 0;
        for (size_t l = 0; l < q.launches.size(); ++l) {
            const VariantKey &k = q.launches[l].key;
            if (q.fused_of[l] < 0) {
                // no launchable launch is left alone while another of its range is there
                if (!fused_launchable(k)) continue;
                for (size_t o = 0; o < q.launches.size(); ++o) {
                    VariantKey a = k, c = q.launches[o].key;
                    const bool same_range = fused_rows(a).lo == fused_rows(c).lo;
                    a.R = c.R = 0;
                    CHECK(o == l || !(a == c) || !same_range, "form 4: launches %zu and %zu unfused", l, o);
                }
                continue;
            }
            const FusedLaunch &F = q.fused[size_t(q.fused_of[l])];
            const RowsRange rr = fused_rows(q.launches[F.leader].key);
            CHECK(k.R >= rr.lo && k.R <= rr.hi, "form 4: launch %zu of R %d in a fused launch of R %d-%d", l, k.R,
                  rr.lo, rr.hi);
            if (!k.local) ++global_members;
        }
        for (const FusedLaunch &F : q.fused) global_groups += q.launches[F.leader].key.local ? 0 : 1;
        // the batch's global pairs span R = 4 to 16: both ranges
        CHECK(global_groups == 2 && global_members >= 4, "form 4: %zu global fused launches of %zu members",
              global_groups, global_members);
        cfg.tuning = &t;
    }
    std::printf("  fused launches: %zu of %zu launches in %zu fused, %zu items\n",
                size_t(std::count_if(p.fused_of.begin(), p.fused_of.end(), [](int32_t x) { return x >= 0; })), nl,
                p.fused.size(), p.fused_items.size());
}

int main()
{
    check_fused_launches();
    check_cost_order();
    check_recentring();
    check_wavefront_tail();
    check_large_gap_extension();
    std::mt19937_64 rng(20261204);
    pa_config dna{};
    dna.gap_open = -10;
    dna.gap_extension = -1;
    dna.match = 4;
    dna.mismatch = -5;
    static const char *kAmino = "ARNDCQEGHILKMFPSTWYVX";
    std::vector<int32_t> matrix(21 * 21);
    for (int i = 0; i < 21; ++i)
        for (int j = 0; j < 21; ++j) matrix[i * 21 + j] = i == j ? 5 : -1 - (i + j) % 3;
    pa_config prot{};
    prot.gap_open = -11;
    prot.gap_extension = -1;
    prot.score_kind = PA_SCORE_MATRIX;
    prot.matrix_dim = 21;
    prot.matrix = matrix.data();
    prot.matrix_symbols = reinterpret_cast<const uint8_t *>(kAmino);

    for (int rep = 0; rep < 2; ++rep) {
        // large enough for several blocks of pairs and several sorted runs per group
        Batch big = random_batch(rng, 150000, "ACGT", 400, false, rep == 1, false);
        check_plan("dna global+local", dna, big);
        pa_config lr = dna;
        lr.last_row_free = 1;
        check_plan("dna free last row", lr, big);
        Batch odd = random_batch(rng, 20000, "ACGTN", 300, false, false, true);
        check_plan("dna with odd bytes", dna, odd);
        Batch pb = random_batch(rng, 40000, kAmino, 500, false, rep == 1, false);
        prot.method = PA_METHOD_LOCAL;
        check_plan("protein local", prot, pb);
        prot.method = PA_METHOD_GLOBAL;
        check_plan("protein per-pair methods", prot, pb);
        check_plan("protein export", prot, pb, nullptr, nullptr, true);
        std::vector<uint8_t> forced(pb.o2.size() - 1, 0);
        std::vector<uint32_t> only;
        for (size_t q = 0; q < forced.size(); q += 3) {
            forced[q] = 32;
            only.push_back(uint32_t(q));
        }
        check_plan("protein int32 subset", prot, pb, forced.data(), &only);
        Batch oq = random_batch(rng, 30000, kAmino, 500, true, false, false);
        prot.method = PA_METHOD_LOCAL;
        check_plan("protein one query", prot, oq);
        Batch odd_p = random_batch(rng, 2000, kAmino, 200, false, false, true);
        check_plan("protein with odd bytes", prot, odd_p);
    }
    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
