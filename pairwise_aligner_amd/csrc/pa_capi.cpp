// pa_capi.cpp -- implementation of the C ABI in include/pa_device.h on top of HIP.
//
// A context owns one device, one stream and a few side streams.  A batch owns the plan (pa_plan.hpp) and its
// device buffers.  A pass enqueues the initial-score kernel and one striped-DP launch per
// kernel variant; when an int16 launch can report saturation, the pass then reads the flags,
// re-plans the flagged pairs at int32 (cached for the next pass) and runs them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "pa_device.h"
#include "pa_kernels.hpp"
#include "pa_learned.hpp"
#include "pa_noalloc.hpp"
#include "pa_plan.hpp"

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;   // false: a slice of a caller's arena (pa_score_batch_noalloc)
    ~DeviceBuffer() { if (ptr && owned) (void)hipFree(ptr); }
};

// Device memory handed out by bumping an offset through one preallocated region instead of
// hipMalloc (pa_score_batch_noalloc's scratch).  Slices are 256-byte aligned.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    size_t refused = 0;   // bytes of the first request that did not fit
};

// hipMalloc, or a slice of `arena` when one is given.
hipError_t dev_alloc(DeviceBuffer &buf, size_t bytes, Arena *arena)
{
    bytes = std::max<size_t>(bytes, 16);
    if (!arena) {
        hipError_t e = hipMalloc(&buf.ptr, bytes);
        if (e == hipSuccess) buf.bytes = bytes;
        return e;
    }
    const size_t off = (arena->used + 255) & ~size_t(255);
    if (off > arena->size || bytes > arena->size - off) {
        if (!arena->refused) arena->refused = bytes;
        return hipErrorOutOfMemory;
    }
    buf.ptr = arena->base + off;
    buf.bytes = bytes;
    buf.owned = false;
    arena->used = off + bytes;
    return hipSuccess;
}

struct PlanDevice {
    pa::Plan plan;
    DeviceBuffer tasks, rowcodes, colcodes, bnd, subset, enc, codes, dp_base;
    DeviceBuffer queues;   // persistent launches: two counters per launch, zero between passes
    DeviceBuffer fused_items;   // wave items of the fused launches (pa::Plan::fused_items)
    std::vector<uint32_t> queues_host;   // their initial zeros (kept until the upload has read them)
    std::vector<pa::CodeTables> code_host;   // per launch (kept until the encoder has read them)
    DeviceBuffer table[pa::kScoreShifts], ptable[pa::kScoreShifts], profile[pa::kScoreShifts];
    std::vector<uint32_t> subset_host;
};

// Kernel variants of one pass run concurrently on kSide side streams (default priority), forked
// from and joined back into the context stream with events, so that their tails overlap.
static constexpr int kSide = 4;

struct pa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kSide] = {};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string last_error;
};

struct pa_batch {
    // The caller's configuration, with the pointers it holds beyond the call made the batch's own:
    // the matrix and symbols copied below, the tuning copied into `tuning` (the int32 re-plan of
    // saturated pairs reads it at a later run), the model cleared once it has been read.
    pa_config cfg{};
    pa_tuning tuning{};
    std::vector<int32_t> matrix;
    std::vector<uint8_t> symbols;
    // What re-planning the saturated pairs needs of the caller's inputs: the offsets, per-pair
    // methods and, for one query, the query bytes (the other residues stay on the device only;
    // the re-plan takes the alphabet from the main plan's residue codes)
    std::vector<uint8_t> query, pair_method;
    std::vector<uint64_t> off1, off2;
    uint64_t n_pairs = 0;
    bool one_query = false;   // seq1 is one query scored against every seq2 (off1 = {0, length})
    bool export_dp = false;   // the kernels also write the final DP rows and columns into dp_out
    DeviceBuffer residues;    // seq1 then seq2 bytes on the device (upload_residues)
    DeviceBuffer dp_out;
    PlanDevice main;
    std::unique_ptr<PlanDevice> recompute;   // int32 re-plan of the pairs that saturated
    std::vector<uint32_t> recompute_set;
    DeviceBuffer pairs, scores, flags;
    bool can_saturate = false;
    Arena *arena = nullptr;                // device buffers from this arena instead of hipMalloc
    hipGraph_t graph = nullptr;            // captured main pass (see launch_main)
    hipGraphExec_t graph_exec = nullptr;
    pa_batch_stats stats{};
    // pa_batch_time: a pair of events around the main launches of every timed pass, read after
    // the last pass, so that timed passes follow each other on the stream without a host wait
    std::vector<hipEvent_t> pass_ev;
    ~pa_batch()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t ev : pass_ev) (void)hipEventDestroy(ev);
    }
};

struct pa_model {
    int device = 0;
    DeviceBuffer params, grad, m1, m2, loss;
    int step = 0;
    float margin = 0.10f;   // relative safety margin of the lane choice (pa_model_set_margin)
};

namespace {

pa_status fail(pa_ctx *ctx, pa_status st, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    return st;
}

pa_status hip_fail(pa_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, e == hipErrorOutOfMemory ? PA_ERR_OUT_OF_MEMORY : PA_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define PA_HIP(call, what)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

template <typename T, typename A>
pa_status upload(pa_ctx *ctx, DeviceBuffer &buf, const std::vector<T, A> &host, Arena *arena = nullptr)
{
    PA_HIP(dev_alloc(buf, host.size() * sizeof(T), arena), "device allocation");
    if (!host.empty())
        PA_HIP(hipMemcpyAsync(buf.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream),
               "hipMemcpyAsync");
    return PA_OK;
}

// Copies the residue bytes of a batch into one device buffer: seq1 bytes [0, end1), then the
// seq2 bytes (the layout the plan's EncodeTask offsets refer to).  Touches nothing of the context
// but its device and stream, so that it can run on a helper thread.
hipError_t copy_residues(int device, hipStream_t stream, const pa::BatchView &v, DeviceBuffer &buf,
                         Arena *arena = nullptr)
{
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    const uint64_t n1 = v.n_pairs ? v.end1() : 0, n2 = v.n_pairs ? v.off2[v.n_pairs] : 0;
    if ((e = dev_alloc(buf, n1 + n2, arena)) != hipSuccess) return e;
    char *dst = static_cast<char *>(buf.ptr);
    if (n1 && (e = hipMemcpyAsync(dst, v.seq1, n1, hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
    if (n2 && (e = hipMemcpyAsync(dst + n1, v.seq2, n2, hipMemcpyHostToDevice, stream)) != hipSuccess) return e;
    return hipSuccess;
}

pa_status upload_residues(pa_ctx *ctx, const pa::BatchView &v, DeviceBuffer &buf)
{
    hipError_t e = copy_residues(ctx->device, ctx->stream, v, buf);
    return e == hipSuccess ? PA_OK : hip_fail(ctx, e, "residue upload");
}

// Uploads a plan and encodes its residue streams on the device from `residues` (upload_residues).
pa_status upload_plan(pa_ctx *ctx, PlanDevice &pd, const DeviceBuffer &residues, Arena *arena = nullptr)
{
    pa_status st;
    const pa::Plan &p = pd.plan;
    if ((st = upload(ctx, pd.tasks, p.tasks, arena)) != PA_OK) return st;
    if ((st = upload(ctx, pd.enc, p.enc, arena)) != PA_OK) return st;
    if ((st = upload(ctx, pd.dp_base, p.dp_base, arena)) != PA_OK) return st;
    PA_HIP(dev_alloc(pd.rowcodes, p.n_rowcodes * 4, arena), "device allocation");
    PA_HIP(dev_alloc(pd.colcodes, p.n_colcodes * 4, arena), "device allocation");
    pd.code_host.resize(p.launches.size());
    for (size_t l = 0; l < p.launches.size(); ++l) pa::code_tables(p, p.launches[l].key, pd.code_host[l]);
    if ((st = upload(ctx, pd.codes, pd.code_host, arena)) != PA_OK) return st;
    for (size_t l = 0; l < p.launches.size(); ++l) {
        const pa::Launch &L = p.launches[l];
        pa::EncodeArgs ea{};
        ea.tasks = static_cast<const pa::Task *>(pd.tasks.ptr);
        ea.enc = static_cast<const pa::EncodeTask *>(pd.enc.ptr);
        ea.task_begin = L.task_begin;
        ea.task_end = L.task_end;
        ea.residues = static_cast<const uint8_t *>(residues.ptr);
        ea.codes = static_cast<const pa::CodeTables *>(pd.codes.ptr) + l;
        ea.rowcodes = static_cast<uint32_t *>(pd.rowcodes.ptr);
        ea.colcodes = static_cast<uint32_t *>(pd.colcodes.ptr);
        ea.stripe_rows = uint32_t(L.key.G) * L.key.R;
        ea.G = L.key.G;
        ea.width = L.key.width;
        ea.score = L.key.score;
        ea.pair_K = p.alphabet;
        hipError_t e = pa::launch_encode(ea, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "encode launch");
    }
    for (int v = 0; v < pa::kScoreShifts; ++v) {
        if ((st = upload(ctx, pd.table[v], pd.plan.table[v], arena)) != PA_OK) return st;
        if ((st = upload(ctx, pd.ptable[v], pd.plan.ptable[v], arena)) != PA_OK) return st;
        if ((st = upload(ctx, pd.profile[v], pd.plan.profile[v], arena)) != PA_OK) return st;
    }
    PA_HIP(dev_alloc(pd.bnd, pd.plan.bnd_entries * 8, arena), "device allocation");
    pd.queues_host.assign(2 * p.launches.size(), 0u);
    if ((st = upload(ctx, pd.queues, pd.queues_host, arena)) != PA_OK) return st;
    if (!p.fused_items.empty())
        if ((st = upload(ctx, pd.fused_items, p.fused_items, arena)) != PA_OK) return st;
    if (!pd.subset_host.empty())
        if ((st = upload(ctx, pd.subset, pd.subset_host, arena)) != PA_OK) return st;
    return PA_OK;
}

pa::KernelArgs kernel_args(const PlanDevice &pd, pa_batch *b, const pa::Launch &L)
{
    const pa::Plan &p = pd.plan;
    pa::KernelArgs a{};
    a.tasks = static_cast<const pa::Task *>(pd.tasks.ptr) + L.task_begin;
    a.ntasks = L.task_end - L.task_begin;
    a.rowcodes = static_cast<const uint32_t *>(pd.rowcodes.ptr);
    a.colcodes = static_cast<const uint32_t *>(pd.colcodes.ptr);
    a.bnd = static_cast<uint32_t *>(pd.bnd.ptr);
    a.scores = static_cast<int32_t *>(b->scores.ptr);
    a.flags = static_cast<uint32_t *>(b->flags.ptr);
    a.error_slot = uint32_t(b->n_pairs);
    // global-method variants read the scores shifted by their drift (pa_cells.hpp)
    const int v = pa::score_shift(L.key);
    if (L.key.score == pa::kScorePairs) {
        a.table = static_cast<const int16_t *>(pd.ptable[v].ptr);
        a.table_entries = uint32_t(p.ptable[v].size());   // dwords
    } else if (L.key.score == pa::kScoreProfile) {
        a.table = static_cast<const int16_t *>(pd.profile[v].ptr);
        a.table_entries = uint32_t(p.profile[v].size());   // dwords
    } else {
        a.table = static_cast<const int16_t *>(pd.table[v].ptr);
        a.table_entries = uint32_t(p.table[v].size());
    }
    a.go = p.go;
    a.ge = p.ge;
    a.lbias = p.lbias;
    a.tbl_lo = p.tbl_lo[v];
    a.tbl_hi = p.tbl_hi[v];
    a.fc = p.fc; a.fr = p.fr; a.lc = p.lc; a.lr = p.lr;
    if ((p.tuning.persistent == 1 || p.tuning.persistent == 2) && L.key.G == 16) {
        const size_t li = size_t(&L - p.launches.data());
        a.queue = static_cast<uint32_t *>(pd.queues.ptr) + 2 * li;
    }
    if (L.key.export_dp) {
        a.dp_out = static_cast<int32_t *>(b->dp_out.ptr);
        a.dp_base = static_cast<const uint64_t *>(pd.dp_base.ptr) + 2ull * L.task_begin;
        a.dp_gap = p.dp_entries;
    }
    return a;
}

// Enqueues the launches of a plan: forked from the context stream onto the plan's streams in
// the plan's order, joined back into the context stream.  self_init: the DP kernels write their
// pairs' initial scores and flags (the main pass); otherwise the initial-score kernel has.
pa_status launch_plan(pa_ctx *ctx, pa_batch *b, PlanDevice &pd, bool self_init)
{
    const pa::Plan &p = pd.plan;
    const size_t nl = p.launches.size();
    int nside = 0;
    for (size_t i = 0; i < nl; ++i) nside = std::max(nside, int(p.launch_stream[i]) + 1);
    const bool fork = nside > 1;
    if (fork) {
        PA_HIP(hipEventRecord(ctx->fork, ctx->stream), "hipEventRecord");
        for (int i = 0; i < nside; ++i) PA_HIP(hipStreamWaitEvent(ctx->side[i], ctx->fork, 0), "hipStreamWaitEvent");
    }
    for (uint32_t li : p.launch_order) {
        const pa::Launch &L = p.launches[li];
        pa::KernelArgs a = kernel_args(pd, b, L);
        a.self_init = self_init ? 1u : 0u;
        const uint32_t lds = pa::lds_bytes(p, L.key);
        hipStream_t st = fork ? ctx->side[p.launch_stream[li]] : ctx->stream;
        const int32_t f = li < p.fused_of.size() ? p.fused_of[li] : -1;
        if (f >= 0) {
            // a fused launch: its leader launches the wave items of every member (the members'
            // keys differ only in R: same tables, same LDS); the other members are not launched
            const pa::FusedLaunch &F = p.fused[size_t(f)];
            if (F.leader != li) continue;
            a.tasks = static_cast<const pa::Task *>(pd.tasks.ptr);
            a.ntasks = 0;
            a.items = static_cast<const uint32_t *>(pd.fused_items.ptr) + F.items_begin;
            a.nitems = F.items_end - F.items_begin;
            a.queue = static_cast<uint32_t *>(pd.queues.ptr) + 2 * size_t(li);
            a.share = float(F.share);
            hipError_t e = pa::launch_fused(L.key, a, lds, st);
            if (e != hipSuccess) return hip_fail(ctx, e, "fused striped DP launch");
            continue;
        }
        hipError_t e = pa::launch_stripe(L.key, a, lds, st);
        if (e != hipSuccess) return hip_fail(ctx, e, "striped DP launch");
    }
    if (fork) {
        for (int i = 0; i < nside; ++i) {
            PA_HIP(hipEventRecord(ctx->join[i], ctx->side[i]), "hipEventRecord");
            PA_HIP(hipStreamWaitEvent(ctx->stream, ctx->join[i], 0), "hipStreamWaitEvent");
        }
    }
    return PA_OK;
}

pa::InitArgs init_args(pa_batch *b, const pa::Plan &p)
{
    pa::InitArgs ia{};
    ia.pairs = static_cast<const pa::PairMeta *>(b->pairs.ptr);
    ia.subset = nullptr;
    ia.n_pairs = b->n_pairs;
    ia.scores = static_cast<int32_t *>(b->scores.ptr);
    ia.flags = static_cast<uint32_t *>(b->flags.ptr);
    ia.go = p.go; ia.ge = p.ge;
    ia.fc = p.fc; ia.fr = p.fr; ia.lc = p.lc; ia.lr = p.lr;
    ia.flag_or = 0;
    return ia;
}

// The main pass: every launch of the main plan, whose kernels initialise their own pairs; the
// initial-score kernel runs first only for batches with pairs that have no task (an empty
// sequence: a closed-form score).  Unless the tuning asks
// for eager launches, a pass with two or more striped launches is captured once into a hipGraph
// and replayed (one launch call per pass, no per-kernel launch gaps, the launches' streams
// overlapping).  A pass with one striped launch runs eagerly: its two kernels run back to back
// either way, and launching them costs less than launching a graph (10,000 DNA pairs of 150 nt:
// 0.089 against 0.096 ms per pass).  The graph holds kernels and event dependencies only.
pa_status enqueue_main(pa_ctx *ctx, pa_batch *b)
{
    if (b->main.plan.n_empty > 0) {
        hipError_t e = pa::launch_init(init_args(b, b->main.plan), ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "init launch");
    }
    return launch_plan(ctx, b, b->main, true);
}

pa_status launch_main(pa_ctx *ctx, pa_batch *b)
{
    if (b->main.plan.tuning.eager || b->main.plan.launches.size() <= 1) return enqueue_main(ctx, b);
    if (!b->graph_exec) {
        PA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        pa_status st = enqueue_main(ctx, b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st != PA_OK) { if (g) (void)hipGraphDestroy(g); return st; }
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        b->graph = g;
        PA_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    PA_HIP(hipGraphLaunch(b->graph_exec, ctx->stream), "hipGraphLaunch");
    return PA_OK;
}

// One pass: the main launches, then, for pairs speculated in int16, the saturation read-back and
// the int32 recomputation.  ev_main (nullable): two events recorded around the main launches.
pa_status run_pass(pa_ctx *ctx, pa_batch *b, const hipEvent_t *ev_main = nullptr)
{
    pa_status st;
    if (ev_main) PA_HIP(hipEventRecord(ev_main[0], ctx->stream), "hipEventRecord");
    if ((st = launch_main(ctx, b)) != PA_OK) return st;
    if (ev_main) PA_HIP(hipEventRecord(ev_main[1], ctx->stream), "hipEventRecord");
    b->stats.n_recomputed = 0;
    if (!b->can_saturate) return PA_OK;

    std::vector<uint32_t> flags(b->n_pairs);
    PA_HIP(hipMemcpyAsync(flags.data(), b->flags.ptr, b->n_pairs * 4, hipMemcpyDeviceToHost, ctx->stream),
           "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    std::vector<uint32_t> sat;
    for (uint64_t p = 0; p < b->n_pairs; ++p)
        if (flags[p] & pa::kFlagSaturated) sat.push_back(uint32_t(p));
    b->stats.n_recomputed = sat.size();
    if (sat.empty()) return PA_OK;

    if (!b->recompute || b->recompute_set != sat) {
        auto pd = std::make_unique<PlanDevice>();
        std::vector<uint8_t> forced(b->n_pairs, 0);
        for (uint32_t p : sat) forced[p] = 32;
        pa::BatchView view{b->one_query ? b->query.data() : nullptr, b->off1.data(), nullptr, b->off2.data(),
                           b->n_pairs, b->pair_method.empty() ? nullptr : b->pair_method.data(), b->one_query,
                           b->main.plan.code};
        std::string err;
        pa_config cfg = b->cfg;
        if ((st = pa::make_plan(cfg, view, forced.data(), &sat, pd->plan, err)) != PA_OK) return fail(ctx, st, err);
        pd->subset_host = sat;
        if ((st = upload_plan(ctx, *pd, b->residues, b->arena)) != PA_OK) return st;
        b->recompute = std::move(pd);
        b->recompute_set = sat;
    }
    // reset the flagged pairs to their initial value, then recompute them in int32 lanes
    pa::InitArgs ia = init_args(b, b->main.plan);
    ia.subset = static_cast<const uint32_t *>(b->recompute->subset.ptr);
    ia.n_pairs = sat.size();
    ia.pairs = static_cast<const pa::PairMeta *>(b->pairs.ptr);
    ia.flag_or = 0x04u;  // PA_PAIR_RECOMPUTED
    // the per-pair width in the main plan says 16; the recomputed pairs report int32
    hipError_t e = pa::launch_init(ia, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "init launch");
    return launch_plan(ctx, b, *b->recompute, false);
}

// ---- learned dispatch helpers ---------------------------------------------------------------

// Feature-kernel arguments for a batch: k-mers over 2-bit codes when the batch uses unitary
// scoring over at most four symbols, else 5-bit residue codes (byte & 31) and k = 3.
void feature_setup(const pa_config &cfg, const pa::BatchView &v, pa::FeatureArgs &fa)
{
    bool seen[256] = {false};
    for (uint64_t i = v.off1[0]; i < v.end1(); ++i) seen[v.seq1[i]] = true;
    for (uint64_t i = v.off2[0]; i < v.off2[v.n_pairs]; ++i) seen[v.seq2[i]] = true;
    int k = 0;
    for (int c = 0; c < 256; ++c) if (seen[c]) ++k;
    const bool small = cfg.score_kind == PA_SCORE_UNITARY && k <= 4;
    int r = 0;
    for (int c = 0; c < 256; ++c) fa.code[c] = small ? uint8_t(seen[c] ? r++ : 0) : uint8_t(c & 31);
    fa.k = small ? 8 : 3;
    fa.bits = small ? 2 : 5;
    fa.smax = float(pa::max_substitution(cfg));
}

pa_status device_features(pa_ctx *ctx, const pa_config &cfg, const pa::BatchView &v, DeviceBuffer &dfeat)
{
    pa::FeatureArgs fa{};
    feature_setup(cfg, v, fa);
    std::vector<uint8_t> s1;
    std::vector<uint64_t> o1(v.n_pairs + 1, 0);
    if (v.one_query) {
        // one query for every pair: uploaded once, read by every pair's wave
        s1.assign(v.s1(0), v.s1(0) + v.len1(0));
        o1.assign({0, v.len1(0)});
        fa.one_query = 1;
    } else {
        s1.assign(v.seq1 + v.off1[0], v.seq1 + v.off1[v.n_pairs]);
        for (uint64_t p = 0; p <= v.n_pairs; ++p) o1[p] = v.off1[p] - v.off1[0];
    }
    std::vector<uint8_t> s2(v.seq2 + v.off2[0], v.seq2 + v.off2[v.n_pairs]);
    std::vector<uint64_t> o2(v.off2, v.off2 + v.n_pairs + 1);
    for (auto &o : o2) o -= v.off2[0];
    DeviceBuffer ds1, ds2, do1, do2;
    pa_status st;
    if ((st = upload(ctx, ds1, s1)) != PA_OK) return st;
    if ((st = upload(ctx, ds2, s2)) != PA_OK) return st;
    if ((st = upload(ctx, do1, o1)) != PA_OK) return st;
    if ((st = upload(ctx, do2, o2)) != PA_OK) return st;
    PA_HIP(hipMalloc(&dfeat.ptr, std::max<size_t>(v.n_pairs * pa::kModelFeatures * 4, 16)), "hipMalloc");
    fa.seq1 = static_cast<const uint8_t *>(ds1.ptr);
    fa.off1 = static_cast<const uint64_t *>(do1.ptr);
    fa.seq2 = static_cast<const uint8_t *>(ds2.ptr);
    fa.off2 = static_cast<const uint64_t *>(do2.ptr);
    fa.n_pairs = v.n_pairs;
    fa.out = static_cast<float *>(dfeat.ptr);
    hipError_t e = pa::launch_features(fa, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "feature launch");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PA_OK;
}

// Lane widths from the learned range predictor: int16 where the bound proves it, or where the
// predicted maximum of H with the model's relative margin + 256 stays below the int16 limit (then the pair is
// tracked and recomputed in int32 if it saturates after all); int32 otherwise.  `cost` receives
// the cost output, in ns (y1 = log2(cost + 1) / 32).
pa_status learned_widths(pa_ctx *ctx, const pa_config &cfg, const pa::BatchView &v, std::vector<uint8_t> &forced,
                         std::vector<float> &cost)
{
    if (!cfg.model) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "PA_LANES_LEARNED needs pa_config.model");
    DeviceBuffer dfeat, dy;
    pa_status st;
    if ((st = device_features(ctx, cfg, v, dfeat)) != PA_OK) return st;
    PA_HIP(hipMalloc(&dy.ptr, std::max<size_t>(v.n_pairs * pa::kModelOutputs * 4, 16)), "hipMalloc");
    hipError_t e = pa::launch_forward(static_cast<const float *>(cfg.model->params.ptr),
                                      static_cast<const float *>(dfeat.ptr), v.n_pairs,
                                      static_cast<float *>(dy.ptr), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "forward launch");
    std::vector<float> y(v.n_pairs * pa::kModelOutputs);
    if (!y.empty())
        PA_HIP(hipMemcpyAsync(y.data(), dy.ptr, y.size() * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const double smax = std::max(1, pa::max_substitution(cfg));
    // the largest score the int16 lanes hold exactly, less a fixed margin
    const double limit = double(pa::int16_score_limit(cfg)) - 256.0;
    forced.assign(v.n_pairs, 0);
    cost.assign(v.n_pairs, 0.0f);
    for (uint64_t p = 0; p < v.n_pairs; ++p) {
        const double mn = double(std::min(v.len1(p), v.off2[p + 1] - v.off2[p]));
        const double pred = std::max(0.0, double(y[p * pa::kModelOutputs])) * smax * mn;
        forced[p] = (pred * (1.0 + double(cfg.model->margin)) < limit) ? 16 : 32;
        cost[p] = float(std::exp2(std::min(40.0, std::max(0.0, double(y[p * pa::kModelOutputs + 1]) * 32.0))) - 1.0);
    }
    return PA_OK;
}

// The wavefront kernel's bounded waits set flags[n_pairs] when they give up; the slot is cleared
// again here (passes do not clear it), so that one failed pass is reported once.
pa_status check_error_slot(pa_ctx *ctx, pa_batch *b)
{
    uint32_t err = 0;
    uint32_t *slot = static_cast<uint32_t *>(b->flags.ptr) + b->n_pairs;
    PA_HIP(hipMemcpy(&err, slot, 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (err) {
        static const uint32_t kZero = 0;
        PA_HIP(hipMemcpy(slot, &kZero, 4, hipMemcpyHostToDevice), "hipMemcpy");
        return fail(ctx, PA_ERR_DEVICE, "wavefront kernel: a boundary wait exceeded its limit");
    }
    return PA_OK;
}

} // namespace

extern "C" {

int pa_abi_version(void) { return PA_ABI_VERSION; }

pa_status pa_ctx_create(pa_ctx **out, int device)
{
    if (!out) return PA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto ctx = std::make_unique<pa_ctx>();
    ctx->device = device;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || device < 0 || device >= count) return PA_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return PA_ERR_DEVICE;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return PA_ERR_DEVICE;
    for (auto &s : ctx->side)
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PA_ERR_DEVICE;
    for (auto &ev : ctx->ev)
        if (hipEventCreate(&ev) != hipSuccess) return PA_ERR_DEVICE;
    if (hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess) return PA_ERR_DEVICE;
    for (auto &ev : ctx->join)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return PA_ERR_DEVICE;
    // the address of this device's copy of pa_score_batch_noalloc's scratch (pa_noalloc.hip)
    if (pa::noalloc_setup(device) != hipSuccess) return PA_ERR_DEVICE;
    *out = ctx.release();
    return PA_OK;
}

pa_status pa_ctx_destroy(pa_ctx *ctx)
{
    if (!ctx) return PA_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(ctx->device);
    for (auto &ev : ctx->ev) if (ev) (void)hipEventDestroy(ev);
    for (auto &ev : ctx->join) if (ev) (void)hipEventDestroy(ev);
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    for (auto &s : ctx->side) if (s) (void)hipStreamDestroy(s);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PA_OK;
}

const char *pa_ctx_last_error(const pa_ctx *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

// Creates a batch of pairs (seq1[p], seq2[p]), or with one_query of (seq1[off1[0], off1[1]),
// seq2[p]) for every p.
static pa_status create_batch(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                              bool one_query, const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs,
                              const uint8_t *pair_method, pa_batch **out, bool export_dp = false,
                              Arena *arena = nullptr)
{
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    auto t0 = std::chrono::steady_clock::now();
    auto b = std::make_unique<pa_batch>();
    b->arena = arena;
    b->cfg = *cfg;
    if (cfg->tuning) {
        b->tuning = *cfg->tuning;
        b->cfg.tuning = &b->tuning;
    }
    if (cfg->score_kind == PA_SCORE_MATRIX && cfg->matrix && cfg->matrix_symbols && cfg->matrix_dim > 0) {
        b->matrix.assign(cfg->matrix, cfg->matrix + size_t(cfg->matrix_dim) * cfg->matrix_dim);
        b->symbols.assign(cfg->matrix_symbols, cfg->matrix_symbols + cfg->matrix_dim);
        b->cfg.matrix = b->matrix.data();
        b->cfg.matrix_symbols = b->symbols.data();
    }
    b->n_pairs = n_pairs;
    b->one_query = one_query;
    const uint64_t n1 = one_query ? 1 : n_pairs;
    b->off1.assign(off1, off1 + n1 + 1);
    b->off2.assign(off2, off2 + n_pairs + 1);
    if (one_query) b->query.assign(seq1, seq1 + off1[1]);
    if (b->query.empty()) b->query.push_back(0);
    if (pair_method) b->pair_method.assign(pair_method, pair_method + n_pairs);

    static const uint8_t kNoResidues[1] = {0};
    pa::BatchView view{seq1 ? seq1 : kNoResidues, b->off1.data(), seq2 ? seq2 : kNoResidues, b->off2.data(), n_pairs,
                       pair_method ? b->pair_method.data() : nullptr, one_query};
    std::string err;
    std::vector<uint8_t> forced;
    std::vector<float> cost;
    pa_status st;
    if (cfg->lane_policy == PA_LANES_LEARNED) {
        if ((st = learned_widths(ctx, b->cfg, view, forced, cost)) != PA_OK) return st;
        if (!pa::resolve_tuning(cfg->tuning).cost_order) cost.clear();
    }
    b->cfg.model = nullptr;   // read above only; the re-plan of saturated pairs forces int32
    b->export_dp = export_dp;
    // The residue bytes go to the device while the host plans (the upload needs nothing from
    // the plan); the helper thread is joined before anything else uses the stream.
    hipError_t upload_err = hipSuccess;
    std::thread uploader([&] { upload_err = copy_residues(ctx->device, ctx->stream, view, b->residues, arena); });
    st = pa::make_plan(b->cfg, view, forced.empty() ? nullptr : forced.data(), nullptr, b->main.plan, err, false,
                       export_dp, cost.empty() ? nullptr : cost.data());
    uploader.join();
    if (upload_err != hipSuccess) return hip_fail(ctx, upload_err, "residue upload");
    if (st != PA_OK) return fail(ctx, st, err);
    const pa::Plan &p = b->main.plan;
    // Only int16 pairs whose bound does not exclude saturation need the flag read-back.
    b->can_saturate = p.n_speculative > 0;

    if ((st = upload_plan(ctx, b->main, b->residues, arena)) != PA_OK) return st;
    if (export_dp) {
        b->dp_out.bytes = std::max<size_t>(p.dp_entries * 8, 16);   // scores, then gap states
        PA_HIP(hipMalloc(&b->dp_out.ptr, b->dp_out.bytes), "hipMalloc");
    }
    if ((st = upload(ctx, b->pairs, p.pairs, arena)) != PA_OK) return st;
    PA_HIP(dev_alloc(b->scores, n_pairs * 4, arena), "device allocation");
    PA_HIP(dev_alloc(b->flags, (n_pairs + 1) * 4, arena), "device allocation");   // + error slot
    // the initial-score kernel clears the flags and the error slot on every pass; a batch without
    // pairs launches nothing, so its error slot starts cleared here (outside any graph capture)
    if (arena) {
        static const uint32_t kZero = 0;
        PA_HIP(hipMemcpyAsync(static_cast<uint32_t *>(b->flags.ptr) + n_pairs, &kZero, 4, hipMemcpyHostToDevice,
                              ctx->stream),
               "hipMemcpyAsync");
    } else {
        PA_HIP(hipMemsetAsync(b->flags.ptr, 0, (n_pairs + 1) * 4, ctx->stream), "hipMemsetAsync");
    }
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");

    b->stats.n_pairs = n_pairs;
    b->stats.cells = p.cells;
    b->stats.padded_cells = p.padded_cells;
    b->stats.n_int16 = p.n16;
    b->stats.n_int32 = p.n32;
    size_t launched = 0;
    for (size_t l = 0; l < p.launches.size(); ++l)
        launched += l >= p.fused_of.size() || p.fused_of[l] < 0 || p.fused[size_t(p.fused_of[l])].leader == l;
    b->stats.n_launches = launched + (p.n_empty > 0 ? 1 : 0);
    b->stats.n_profiled = p.n_profiled;
    b->stats.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = b.release();
    return PA_OK;
}

pa_status pa_batch_create(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                          const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs,
                          const uint8_t *pair_method, pa_batch **out)
{
    if (!ctx || !cfg || !out || !off1 || !off2 || (n_pairs && (!seq1 || !seq2)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    return create_batch(ctx, cfg, seq1, off1, false, seq2, off2, n_pairs, pair_method, out);
}

pa_status pa_batch_create_one_to_many(pa_ctx *ctx, const pa_config *cfg, const uint8_t *query, uint64_t query_len,
                                      const uint8_t *seq2, const uint64_t *off2, uint64_t n_targets,
                                      const uint8_t *pair_method, pa_batch **out)
{
    if (!ctx || !cfg || !out || !off2 || (query_len && !query) || (n_targets && !seq2))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    const uint64_t off1[2] = {0, query_len};
    return create_batch(ctx, cfg, query, off1, true, seq2, off2, n_targets, pair_method, out);
}

pa_status pa_batch_run(pa_ctx *ctx, pa_batch *b)
{
    if (!ctx || !b) return PA_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    pa_status st = run_pass(ctx, b);
    if (st != PA_OK) return st;
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return check_error_slot(ctx, b);
}

pa_status pa_batch_results(pa_ctx *ctx, pa_batch *b, int32_t *scores, uint8_t *flags)
{
    if (!ctx || !b || (!scores && b->n_pairs)) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (b->n_pairs == 0) return PA_OK;
    PA_HIP(hipMemcpyAsync(scores, b->scores.ptr, b->n_pairs * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    std::vector<uint32_t> f(b->n_pairs);
    PA_HIP(hipMemcpyAsync(f.data(), b->flags.ptr, b->n_pairs * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (flags) {
        for (uint64_t p = 0; p < b->n_pairs; ++p) {
            uint32_t v = f[p];
            if (v & 0x04u) v = (v & ~0x01u) | 0x02u;   // recomputed pairs end in int32 lanes
            flags[p] = uint8_t(v & 0xFFu);
        }
    }
    return PA_OK;
}

pa_status pa_batch_stats_get(const pa_batch *b, pa_batch_stats *out)
{
    if (!b || !out) return PA_ERR_INVALID_ARGUMENT;
    *out = b->stats;
    return PA_OK;
}

pa_status pa_batch_time(pa_ctx *ctx, pa_batch *b, int iters, double *ms_total, double *ms_main)
{
    if (!ctx || !b || iters < 1) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "invalid argument");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    if (ms_main) {
        while (b->pass_ev.size() < 2 * size_t(iters)) {
            hipEvent_t ev;
            PA_HIP(hipEventCreate(&ev), "hipEventCreate");
            b->pass_ev.push_back(ev);
        }
    }
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    PA_HIP(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    // the passes follow each other on the stream (the host waits only where a pass reads its
    // saturation flags back); the main launches' events are read once all passes have ended
    for (int i = 0; i < iters; ++i) {
        pa_status st = run_pass(ctx, b, ms_main ? &b->pass_ev[2 * size_t(i)] : nullptr);
        if (st != PA_OK) return st;
    }
    PA_HIP(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
    PA_HIP(hipEventSynchronize(ctx->ev[1]), "hipEventSynchronize");
    float ms = 0;
    PA_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
    if (ms_total) *ms_total = ms;
    if (ms_main) {
        double sum = 0;
        for (int i = 0; i < iters; ++i) {
            float m = 0;
            PA_HIP(hipEventElapsedTime(&m, b->pass_ev[2 * size_t(i)], b->pass_ev[2 * size_t(i) + 1]),
                   "hipEventElapsedTime");
            sum += m;
        }
        *ms_main = sum / iters;
    }
    return check_error_slot(ctx, b);
}

pa_status pa_batch_destroy(pa_ctx *ctx, pa_batch *b)
{
    if (!b) return PA_ERR_INVALID_ARGUMENT;
    if (ctx) (void)hipSetDevice(ctx->device);
    delete b;
    return PA_OK;
}

int pa_model_param_count(void) { return pa::kModelParams; }

pa_status pa_model_create(pa_ctx *ctx, uint64_t seed, pa_model **out)
{
    if (!ctx || !out) return PA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto m = std::make_unique<pa_model>();
    m->device = ctx->device;
    // He initialisation with a splitmix64 + Box-Muller generator; biases zero
    std::vector<float> p(pa::kModelParams, 0.0f);
    uint64_t state = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return double((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto gauss = [&]() { return std::sqrt(-2.0 * std::log(next() + 1e-300)) * std::cos(6.283185307179586 * next()); };
    const int F = pa::kModelFeatures, H = pa::kModelHidden, O = pa::kModelOutputs;
    int off = 0;
    for (int i = 0; i < F * H; ++i) p[off + i] = float(gauss() * std::sqrt(2.0 / F));
    off += F * H + H;
    for (int i = 0; i < H * H; ++i) p[off + i] = float(gauss() * std::sqrt(2.0 / H));
    off += H * H + H;
    for (int i = 0; i < H * O; ++i) p[off + i] = float(gauss() * std::sqrt(1.0 / H));
    pa_status st;
    if ((st = upload(ctx, m->params, p)) != PA_OK) return st;
    std::vector<float> zeros(pa::kModelParams, 0.0f);
    if ((st = upload(ctx, m->grad, zeros)) != PA_OK) return st;
    if ((st = upload(ctx, m->m1, zeros)) != PA_OK) return st;
    if ((st = upload(ctx, m->m2, zeros)) != PA_OK) return st;
    std::vector<float> one(4, 0.0f);
    if ((st = upload(ctx, m->loss, one)) != PA_OK) return st;
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    *out = m.release();
    return PA_OK;
}

pa_status pa_model_destroy(pa_model *model)
{
    if (!model) return PA_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(model->device);
    delete model;
    return PA_OK;
}

pa_status pa_model_set_params(pa_ctx *ctx, pa_model *model, const float *params)
{
    if (!ctx || !model || !params) return PA_ERR_INVALID_ARGUMENT;
    PA_HIP(hipMemcpy(model->params.ptr, params, pa::kModelParams * 4, hipMemcpyHostToDevice), "hipMemcpy");
    std::vector<float> zeros(pa::kModelParams, 0.0f);
    PA_HIP(hipMemcpy(model->m1.ptr, zeros.data(), pa::kModelParams * 4, hipMemcpyHostToDevice), "hipMemcpy");
    PA_HIP(hipMemcpy(model->m2.ptr, zeros.data(), pa::kModelParams * 4, hipMemcpyHostToDevice), "hipMemcpy");
    model->step = 0;
    return PA_OK;
}

pa_status pa_model_set_margin(pa_model *model, float margin)
{
    if (!model || !(margin >= 0.0f && margin <= 10.0f)) return PA_ERR_INVALID_ARGUMENT;
    model->margin = margin;
    return PA_OK;
}

pa_status pa_model_get_params(pa_ctx *ctx, const pa_model *model, float *params)
{
    if (!ctx || !model || !params) return PA_ERR_INVALID_ARGUMENT;
    PA_HIP(hipMemcpy(params, model->params.ptr, pa::kModelParams * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return PA_OK;
}

pa_status pa_features(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                      const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, float *features)
{
    if (!ctx || !cfg || !off1 || !off2 || !features || (n_pairs && (!seq1 || !seq2)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (n_pairs == 0) return PA_OK;
    pa::BatchView v{seq1, off1, seq2, off2, n_pairs, nullptr};
    DeviceBuffer dfeat;
    pa_status st = device_features(ctx, *cfg, v, dfeat);
    if (st != PA_OK) return st;
    PA_HIP(hipMemcpy(features, dfeat.ptr, n_pairs * pa::kModelFeatures * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return PA_OK;
}

pa_status pa_model_forward(pa_ctx *ctx, const pa_model *model, const float *x, uint64_t n, float *y)
{
    if (!ctx || !model || (n && (!x || !y))) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return PA_OK;
    DeviceBuffer dx, dy;
    std::vector<float> hx(x, x + n * pa::kModelFeatures);
    pa_status st;
    if ((st = upload(ctx, dx, hx)) != PA_OK) return st;
    PA_HIP(hipMalloc(&dy.ptr, n * pa::kModelOutputs * 4), "hipMalloc");
    hipError_t e = pa::launch_forward(static_cast<const float *>(model->params.ptr), static_cast<const float *>(dx.ptr),
                                      n, static_cast<float *>(dy.ptr), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "forward launch");
    PA_HIP(hipMemcpyAsync(y, dy.ptr, n * pa::kModelOutputs * 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PA_OK;
}

static pa_status loss_grad_device(pa_ctx *ctx, const pa_model *model, const float *x, const float *target,
                                  const float *weight, uint64_t n, DeviceBuffer &dx, DeviceBuffer &dt, DeviceBuffer &dw)
{
    std::vector<float> hx(x, x + n * pa::kModelFeatures), ht(target, target + n * pa::kModelOutputs);
    pa_status st;
    if ((st = upload(ctx, dx, hx)) != PA_OK) return st;
    if ((st = upload(ctx, dt, ht)) != PA_OK) return st;
    std::vector<float> hw;
    if (weight) {
        hw.assign(weight, weight + n * pa::kModelOutputs);
        if ((st = upload(ctx, dw, hw)) != PA_OK) return st;
    }
    hipError_t e = pa::launch_loss_grad(static_cast<const float *>(model->params.ptr),
                                        static_cast<const float *>(dx.ptr), static_cast<const float *>(dt.ptr),
                                        weight ? static_cast<const float *>(dw.ptr) : nullptr, n,
                                        static_cast<float *>(model->grad.ptr), static_cast<float *>(model->loss.ptr),
                                        ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "loss/gradient launch");
    // the uploads above read host vectors that end with this function
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PA_OK;
}

pa_status pa_model_loss_grad(pa_ctx *ctx, const pa_model *model, const float *x, const float *target,
                             const float *weight, uint64_t n, float *loss, float *grad)
{
    if (!ctx || !model || !x || !target || n == 0) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "invalid argument");
    DeviceBuffer dx, dt, dw;
    pa_status st = loss_grad_device(ctx, model, x, target, weight, n, dx, dt, dw);
    if (st != PA_OK) return st;
    if (grad)
        PA_HIP(hipMemcpyAsync(grad, model->grad.ptr, pa::kModelParams * 4, hipMemcpyDeviceToHost, ctx->stream),
               "hipMemcpyAsync");
    if (loss) PA_HIP(hipMemcpyAsync(loss, model->loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PA_OK;
}

pa_status pa_model_train_step(pa_ctx *ctx, pa_model *model, const float *x, const float *target, const float *weight,
                              uint64_t n, float learning_rate, float *loss)
{
    if (!ctx || !model || !x || !target || n == 0) return fail(ctx, PA_ERR_INVALID_ARGUMENT, "invalid argument");
    DeviceBuffer dx, dt, dw;
    pa_status st = loss_grad_device(ctx, model, x, target, weight, n, dx, dt, dw);
    if (st != PA_OK) return st;
    model->step += 1;
    hipError_t e = pa::launch_adam(static_cast<float *>(model->params.ptr), static_cast<const float *>(model->grad.ptr),
                                   static_cast<float *>(model->m1.ptr), static_cast<float *>(model->m2.ptr),
                                   learning_rate, model->step, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "adam launch");
    if (loss) PA_HIP(hipMemcpyAsync(loss, model->loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PA_OK;
}

pa_status pa_dp_labels(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                       const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, int32_t *max_h, float *cost)
{
    if (!ctx || !cfg || !off1 || !off2 || !max_h || (n_pairs && (!seq1 || !seq2)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (n_pairs == 0) return PA_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    // every pair in int32 lanes with the running maximum kept; launches run one after the other
    // so that each one's HIP-event time can be shared out over its tasks by modelled cost
    pa_batch b{};
    b.cfg = *cfg;
    b.n_pairs = n_pairs;
    pa::BatchView view{seq1, off1, seq2, off2, n_pairs, nullptr};
    std::vector<uint8_t> forced(n_pairs, 32);
    std::string err;
    pa_status st = pa::make_plan(*cfg, view, forced.data(), nullptr, b.main.plan, err, true);
    if (st != PA_OK) return fail(ctx, st, err);
    if ((st = upload_residues(ctx, view, b.residues)) != PA_OK) return st;
    if ((st = upload_plan(ctx, b.main, b.residues)) != PA_OK) return st;
    if ((st = upload(ctx, b.pairs, b.main.plan.pairs)) != PA_OK) return st;
    PA_HIP(hipMalloc(&b.scores.ptr, n_pairs * 4), "hipMalloc");
    PA_HIP(hipMalloc(&b.flags.ptr, (n_pairs + 1) * 4), "hipMalloc");
    DeviceBuffer dmax;
    std::vector<int32_t> init(n_pairs, INT32_MIN);
    if ((st = upload(ctx, dmax, init)) != PA_OK) return st;
    hipError_t e = pa::launch_init(init_args(&b, b.main.plan), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "init launch");
    std::vector<double> pair_cost(n_pairs, 0.0);
    for (const pa::Launch &L : b.main.plan.launches) {
        pa::KernelArgs a = kernel_args(b.main, &b, L);
        a.maxh = static_cast<int32_t *>(dmax.ptr);
        const uint32_t lds = pa::lds_bytes(b.main.plan, L.key);
        PA_HIP(hipEventRecord(ctx->ev[2], ctx->stream), "hipEventRecord");
        e = pa::launch_stripe(L.key, a, lds, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "striped DP launch");
        PA_HIP(hipEventRecord(ctx->ev[3], ctx->stream), "hipEventRecord");
        PA_HIP(hipEventSynchronize(ctx->ev[3]), "hipEventSynchronize");
        float ms = 0;
        PA_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "hipEventElapsedTime");
        double total = 0;
        for (uint32_t t = L.task_begin; t < L.task_end; ++t) total += pa::task_cost(L.key, b.main.plan.tasks[t], b.main.plan.tuning);
        for (uint32_t t = L.task_begin; t < L.task_end; ++t) {
            const pa::Task &tk = b.main.plan.tasks[t];
            const double share = double(ms) * 1e6 * pa::task_cost(L.key, tk, b.main.plan.tuning) / std::max(total, 1e-9);
            for (int q = 0; q < 2; ++q)
                if (tk.pair[q] != pa::kNone) pair_cost[tk.pair[q]] = share / (tk.pair[1] != pa::kNone ? 2.0 : 1.0);
        }
    }
    PA_HIP(hipMemcpy(max_h, dmax.ptr, n_pairs * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if ((st = check_error_slot(ctx, &b)) != PA_OK) return st;
    if (cost)
        for (uint64_t p = 0; p < n_pairs; ++p) cost[p] = float(pair_cost[p]);   // nanoseconds
    return PA_OK;
}

pa_status pa_tune(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                  const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, const uint8_t *pair_method,
                  uint64_t sample_pairs, int iters, pa_tuning *out)
{
    if (!ctx || !cfg || !off1 || !off2 || !out || (n_pairs && (!seq1 || !seq2)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (iters <= 0) iters = 3;
    // the sample: every k-th pair, which keeps the batch's length distribution
    if (sample_pairs == 0) sample_pairs = std::max<uint64_t>(2000, n_pairs / 20);
    const uint64_t k = std::max<uint64_t>(1, n_pairs / std::max<uint64_t>(sample_pairs, 1));
    std::vector<uint8_t> s1, s2, pm;
    std::vector<uint64_t> o1{0}, o2{0};
    for (uint64_t p = 0; p < n_pairs; p += k) {
        s1.insert(s1.end(), seq1 + off1[p], seq1 + off1[p + 1]);
        s2.insert(s2.end(), seq2 + off2[p], seq2 + off2[p + 1]);
        o1.push_back(s1.size());
        o2.push_back(s2.size());
        if (pair_method) pm.push_back(pair_method[p]);
    }
    const uint64_t ns = o1.size() - 1;
    if (s1.empty()) s1.push_back(0);
    if (s2.empty()) s2.push_back(0);

    pa::Tuning base = pa::resolve_tuning(cfg->tuning);
    auto to_c = [](const pa::Tuning &t) {
        pa_tuning c{};
        c.cell_cost = float(t.cell_cost);
        c.step_cost = float(t.step_cost);
        c.rows_mask = t.rows_mask;
        c.wavefront_min_rows = t.wavefront_min_rows;
        c.streams = t.streams;
        c.eager = t.eager ? 1 : 0;
        c.cost_order = t.cost_order ? 1 : 0;
        c.persistent = t.persistent;
        c.wavefront_tail = t.wavefront_tail > 0 ? float(t.wavefront_tail) : -1.0f;
        return c;
    };
    // time of `iters` passes with tuning t (after one untimed pass), on the sample or the batch
    auto measure = [&](const pa_tuning &t, bool full, double &ms) -> pa_status {
        pa_config c = *cfg;
        c.tuning = &t;
        pa_batch *b = nullptr;
        pa_status st = full ? pa_batch_create(ctx, &c, seq1, off1, seq2, off2, n_pairs, pair_method, &b)
                            : pa_batch_create(ctx, &c, s1.data(), o1.data(), s2.data(), o2.data(), ns,
                                              pm.empty() ? nullptr : pm.data(), &b);
        if (st != PA_OK) return st;
        st = pa_batch_run(ctx, b);
        double total = 0;
        if (st == PA_OK) st = pa_batch_time(ctx, b, iters, &total, nullptr);
        pa_batch_destroy(ctx, b);
        ms = total / iters;
        return st;
    };
    pa_status st;

    // Stage 1, on the sample: the shape model (tile height R per pair and its allowed values).
    // These decide per-pair efficiency, which a sample with the batch's length mix shows.
    pa_tuning best = to_c(base);
    double best_ms = 0;
    if ((st = measure(best, false, best_ms)) != PA_OK) return st;
    const float step[] = {8, 16, 24, 40, 64};
    const uint32_t rows[] = {0x7F, 0x55, 0x2A, 0x7E, 0x3F, 0x78, 0x1E};
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int param = 0; param < 2; ++param) {
            const int nv = param == 0 ? 5 : 7;
            for (int v = 0; v < nv; ++v) {
                pa_tuning t = best;
                if (param == 0) t.step_cost = step[v]; else t.rows_mask = rows[v];
                if (std::memcmp(&t, &best, sizeof t) == 0) continue;
                double ms = 0;
                if ((st = measure(t, false, ms)) != PA_OK) return st;
                if (ms < best_ms * 0.98) { best = t; best_ms = ms; }
            }
        }
    }

    // Stage 2, on the whole batch: the launch-level parameters, whose effect depends on the
    // batch's total work against its longest tasks (the wavefront threshold, streams, graph
    // replay), and the stage-1 choice against the defaults.
    const pa_tuning start = to_c(base);
    double full_best = 0;
    if ((st = measure(start, true, full_best)) != PA_OK) return st;
    pa_tuning full = start;
    if (std::memcmp(&best, &start, sizeof best) != 0) {
        double ms = 0;
        if ((st = measure(best, true, ms)) != PA_OK) return st;
        if (ms < full_best) { full = best; full_best = ms; }
    }
    const uint32_t wf[] = {2048, 3072, 4096};
    const uint32_t streams[] = {2, 4};
    const uint32_t eager[] = {0, 1};
    const uint32_t persistent[] = {0, 1, 2, 3, 4};
    const float tails[] = {-1.0f, 0.5f, 1.0f};
    for (int param = 0; param < 5; ++param) {
        const int nv = param == 3 ? 5 : (param == 0 || param == 4 ? 3 : 2);
        for (int v = 0; v < nv; ++v) {
            pa_tuning t = full;
            if (param == 0) t.wavefront_min_rows = wf[v];
            else if (param == 1) t.streams = streams[v];
            else if (param == 2) t.eager = eager[v];
            else if (param == 3) t.persistent = persistent[v];
            else t.wavefront_tail = tails[v];
            if (std::memcmp(&t, &full, sizeof t) == 0) continue;
            double ms = 0;
            if ((st = measure(t, true, ms)) != PA_OK) return st;
            if (ms < full_best * 0.98) { full = t; full_best = ms; }
        }
    }
    // The pick was accepted on single timings a few per cent apart, where one lucky timing can
    // carry a slower candidate: time it against the starting parameters twice more, alternately,
    // and keep it only if it is faster over both rounds.
    if (std::memcmp(&full, &start, sizeof full) != 0) {
        double sum_full = 0, sum_start = 0;
        for (int round = 0; round < 2; ++round) {
            double ms = 0;
            if ((st = measure(start, true, ms)) != PA_OK) return st;
            sum_start += ms;
            if ((st = measure(full, true, ms)) != PA_OK) return st;
            sum_full += ms;
        }
        if (sum_full >= sum_start) full = start;
    }
    best = full;
    *out = best;
    return PA_OK;
}

pa_status pa_score_batch(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                         const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, const uint8_t *pair_method,
                         int32_t *scores, uint8_t *flags)
{
    pa_batch *b = nullptr;
    pa_status st = pa_batch_create(ctx, cfg, seq1, off1, seq2, off2, n_pairs, pair_method, &b);
    if (st != PA_OK) return st;
    st = pa_batch_run(ctx, b);
    if (st == PA_OK) st = pa_batch_results(ctx, b, scores, flags);
    pa_batch_destroy(ctx, b);
    return st;
}

pa_status pa_score_batch_noalloc(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                                 const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs,
                                 const uint8_t *pair_method, int32_t *scores, uint8_t *flags)
{
    if (!ctx || !cfg || !off1 || !off2 || (n_pairs && (!seq1 || !seq2 || !scores)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (cfg->lane_policy == PA_LANES_LEARNED)
        return fail(ctx, PA_ERR_UNSUPPORTED, "pa_score_batch_noalloc: PA_LANES_LEARNED allocates its features; "
                                             "use another lane policy");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    void *base = pa::noalloc_scratch(ctx->device);
    if (!base) return fail(ctx, PA_ERR_DEVICE, "pa_score_batch_noalloc: no scratch on this device");
    // calls on one device share its scratch: one at a time
    std::lock_guard<std::mutex> lock(pa::noalloc_call_mutex(ctx->device));
    Arena arena{static_cast<char *>(base), pa::kNoallocScratchBytes, 0, 0};
    // eager launches: instantiating a graph is left to callers that may allocate
    pa_config c = *cfg;
    pa_tuning t = cfg->tuning ? *cfg->tuning : pa_tuning{};
    t.eager = 1;
    c.tuning = &t;
    auto too_big = [&](pa_status st) {
        if (st == PA_ERR_OUT_OF_MEMORY && arena.refused)
            return fail(ctx, st, "pa_score_batch_noalloc: the batch needs more than the " +
                                     std::to_string(pa::kNoallocScratchBytes >> 20) + " MiB scratch (" +
                                     std::to_string(arena.used) + " bytes in use, " + std::to_string(arena.refused) +
                                     " more requested)");
        return st;
    };
    pa_batch *b = nullptr;
    pa_status st = create_batch(ctx, &c, seq1, off1, false, seq2, off2, n_pairs, pair_method, &b, false, &arena);
    if (st != PA_OK) return too_big(st);
    std::unique_ptr<pa_batch> hold(b);
    if ((st = run_pass(ctx, b)) != PA_OK) return too_big(st);
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if ((st = check_error_slot(ctx, b)) != PA_OK) return st;
    return pa_batch_results(ctx, b, scores, flags);
}

pa_status pa_score_batch_dp(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                            const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, const uint8_t *pair_method,
                            int32_t *scores, uint8_t *flags, int32_t *dp_column, int32_t *dp_row)
{
    return pa_score_batch_dp_cells(ctx, cfg, seq1, off1, seq2, off2, n_pairs, pair_method, scores, flags, dp_column,
                                   dp_row, nullptr, nullptr);
}

pa_status pa_score_batch_dp_cells(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                                  const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs,
                                  const uint8_t *pair_method, int32_t *scores, uint8_t *flags, int32_t *dp_column,
                                  int32_t *dp_row, int32_t *dp_column_gap, int32_t *dp_row_gap)
{
    if (!ctx || !cfg || !off1 || !off2 || (n_pairs && (!seq1 || !seq2 || !dp_column || !dp_row)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    if (!dp_column_gap != !dp_row_gap)
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "dp_column_gap and dp_row_gap go together");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    pa_batch *b = nullptr;
    pa_status st = create_batch(ctx, cfg, seq1, off1, false, seq2, off2, n_pairs, pair_method, &b, true);
    if (st != PA_OK) return st;
    std::unique_ptr<pa_batch> hold(b);
    if ((st = pa_batch_run(ctx, b)) != PA_OK) return st;
    if ((st = pa_batch_results(ctx, b, scores, flags)) != PA_OK) return st;
    const uint64_t n1 = off1[n_pairs] - off1[0] + n_pairs, n2 = off2[n_pairs] - off2[0] + n_pairs;
    const bool gaps = dp_column_gap != nullptr;
    if (n_pairs) {
        // the device buffer holds the columns, then the rows (n1 + n2 entries), then the gap halves
        const char *dev = static_cast<const char *>(b->dp_out.ptr);
        PA_HIP(hipMemcpy(dp_column, dev, n1 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        PA_HIP(hipMemcpy(dp_row, dev + n1 * 4, n2 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        if (gaps) {
            PA_HIP(hipMemcpy(dp_column_gap, dev + (n1 + n2) * 4, n1 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            PA_HIP(hipMemcpy(dp_row_gap, dev + (2 * n1 + n2) * 4, n2 * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        }
    }
    // pairs with an empty sequence have no task: the vector along the other sequence keeps its
    // initial cells (H, H + go + ge) (affine_initialisation_strategy.hpp:38-51; local: every end
    // free), and the vector of one cell holds (H(0, n) or H(m, 0), go + ge)
    const bool local_all = cfg->method == PA_METHOD_LOCAL;
    const int32_t gogo = int32_t(cfg->gap_open + cfg->gap_extension);
    for (uint64_t p = 0; p < n_pairs; ++p) {
        const uint64_t m = off1[p + 1] - off1[p], n = off2[p + 1] - off2[p];
        if (m && n) continue;
        const bool local = pair_method ? pair_method[p] == PA_METHOD_LOCAL : local_all;
        const bool fc = local || cfg->first_column_free, fr = local || cfg->first_row_free;
        auto h = [&](uint64_t k, bool free_end) -> int32_t {
            return (k == 0 || free_end) ? 0 : int32_t(cfg->gap_open + int64_t(cfg->gap_extension) * int64_t(k));
        };
        const uint64_t c0 = (off1[p] - off1[0]) + p, r0 = (off2[p] - off2[0]) + p;
        int32_t *col = dp_column + c0, *row = dp_row + r0;
        if (m == 0) {
            for (uint64_t j = 0; j <= n; ++j) row[j] = h(j, fr);
            col[0] = h(n, fr);
        } else {
            for (uint64_t i = 0; i <= m; ++i) col[i] = h(i, fc);
            row[0] = h(m, fc);
        }
        if (gaps) {
            int32_t *cg = dp_column_gap + c0, *rg = dp_row_gap + r0;
            if (m == 0) {
                for (uint64_t j = 0; j <= n; ++j) rg[j] = h(j, fr) + gogo;
                cg[0] = gogo;
            } else {
                for (uint64_t i = 0; i <= m; ++i) cg[i] = h(i, fc) + gogo;
                rg[0] = gogo;
            }
        }
    }
    return PA_OK;
}

pa_status pa_score_one_to_many(pa_ctx *ctx, const pa_config *cfg, const uint8_t *query, uint64_t query_len,
                               const uint8_t *seq2, const uint64_t *off2, uint64_t n_targets,
                               const uint8_t *pair_method, int32_t *scores, uint8_t *flags)
{
    pa_batch *b = nullptr;
    pa_status st = pa_batch_create_one_to_many(ctx, cfg, query, query_len, seq2, off2, n_targets, pair_method, &b);
    if (st != PA_OK) return st;
    st = pa_batch_run(ctx, b);
    if (st == PA_OK) st = pa_batch_results(ctx, b, scores, flags);
    pa_batch_destroy(ctx, b);
    return st;
}

} // extern "C"
