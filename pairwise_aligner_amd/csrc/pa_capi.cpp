// pa_capi.cpp -- implementation of the C ABI in include/pa_device.h on top of HIP.
//
// A context owns one device, one stream and a few side streams.  A batch owns the plan (pa_plan.hpp) and its
// device buffers.  A pass enqueues the initial-score kernel and one striped-DP launch per
// kernel variant; when an int16 launch can report saturation, the pass then reads the flags,
// re-plans the flagged pairs at int32 (cached for the next pass) and runs them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "pa_device.h"
#include "pa_kernels.hpp"
#include "pa_plan.hpp"

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~DeviceBuffer() { if (ptr) (void)hipFree(ptr); }
};

struct PlanDevice {
    pa::Plan plan;
    DeviceBuffer tasks, rowcodes, colcodes, bnd, table, subset;
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
    pa_config cfg{};
    std::vector<int32_t> matrix;
    std::vector<uint8_t> symbols;
    // caller's inputs (host copies, kept for re-planning saturated pairs)
    std::vector<uint8_t> seq1, seq2, pair_method;
    std::vector<uint64_t> off1, off2;
    uint64_t n_pairs = 0;
    PlanDevice main;
    std::unique_ptr<PlanDevice> recompute;   // int32 re-plan of the pairs that saturated
    std::vector<uint32_t> recompute_set;
    DeviceBuffer pairs, scores, flags;
    bool can_saturate = false;
    pa_batch_stats stats{};
    double main_ms_accum = 0;
    uint64_t main_ms_count = 0;
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

template <typename T>
pa_status upload(pa_ctx *ctx, DeviceBuffer &buf, const std::vector<T> &host)
{
    const size_t bytes = std::max<size_t>(host.size() * sizeof(T), 16);
    PA_HIP(hipMalloc(&buf.ptr, bytes), "hipMalloc");
    buf.bytes = bytes;
    if (!host.empty())
        PA_HIP(hipMemcpyAsync(buf.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream),
               "hipMemcpyAsync");
    return PA_OK;
}

pa_status upload_plan(pa_ctx *ctx, PlanDevice &pd)
{
    pa_status st;
    if ((st = upload(ctx, pd.tasks, pd.plan.tasks)) != PA_OK) return st;
    if ((st = upload(ctx, pd.rowcodes, pd.plan.rowcodes)) != PA_OK) return st;
    if ((st = upload(ctx, pd.colcodes, pd.plan.colcodes)) != PA_OK) return st;
    if ((st = upload(ctx, pd.table, pd.plan.table)) != PA_OK) return st;
    const size_t bnd_bytes = std::max<size_t>(pd.plan.bnd_entries * 8, 16);
    PA_HIP(hipMalloc(&pd.bnd.ptr, bnd_bytes), "hipMalloc");
    pd.bnd.bytes = bnd_bytes;
    if (!pd.subset_host.empty())
        if ((st = upload(ctx, pd.subset, pd.subset_host)) != PA_OK) return st;
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
    a.table = static_cast<const int16_t *>(pd.table.ptr);
    a.table_entries = uint32_t(p.table.size());
    a.go = p.go;
    a.ge = p.ge;
    a.tbl_lo = p.tbl_lo;
    a.tbl_hi = p.tbl_hi;
    a.fc = p.fc; a.fr = p.fr; a.lc = p.lc; a.lr = p.lr;
    return a;
}

pa_status launch_plan(pa_ctx *ctx, pa_batch *b, PlanDevice &pd, bool timed)
{
    if (timed) PA_HIP(hipEventRecord(ctx->ev[2], ctx->stream), "hipEventRecord");
    const size_t nl = pd.plan.launches.size();
    const int nside = int(std::min<size_t>(nl, kSide));
    if (nside > 1) {
        PA_HIP(hipEventRecord(ctx->fork, ctx->stream), "hipEventRecord");
        for (int i = 0; i < nside; ++i) PA_HIP(hipStreamWaitEvent(ctx->side[i], ctx->fork, 0), "hipStreamWaitEvent");
    }
    // launches are ordered by decreasing work: the largest start first
    std::vector<size_t> order(nl);
    for (size_t i = 0; i < nl; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return pd.plan.launches[x].padded_cells > pd.plan.launches[y].padded_cells;
    });
    for (size_t i = 0; i < nl; ++i) {
        const pa::Launch &L = pd.plan.launches[order[i]];
        const pa::KernelArgs a = kernel_args(pd, b, L);
        const uint32_t lds = pd.plan.table.empty() ? 0 : pa::lds_table_bytes(pd.plan.alphabet, pd.plan.stride);
        hipStream_t st = nside > 1 ? ctx->side[i % nside] : ctx->stream;
        hipError_t e = pa::launch_stripe(L.key, a, lds, st);
        if (e != hipSuccess) return hip_fail(ctx, e, "striped DP launch");
    }
    if (nside > 1) {
        for (int i = 0; i < nside; ++i) {
            PA_HIP(hipEventRecord(ctx->join[i], ctx->side[i]), "hipEventRecord");
            PA_HIP(hipStreamWaitEvent(ctx->stream, ctx->join[i], 0), "hipStreamWaitEvent");
        }
    }
    if (timed) {
        PA_HIP(hipEventRecord(ctx->ev[3], ctx->stream), "hipEventRecord");
        PA_HIP(hipEventSynchronize(ctx->ev[3]), "hipEventSynchronize");
        float ms = 0;
        PA_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]), "hipEventElapsedTime");
        b->main_ms_accum += ms;
        b->main_ms_count += 1;
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

pa_status run_pass(pa_ctx *ctx, pa_batch *b, bool timed)
{
    pa_status st;
    hipError_t e = pa::launch_init(init_args(b, b->main.plan), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "init launch");
    if ((st = launch_plan(ctx, b, b->main, timed)) != PA_OK) return st;
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
        pa::BatchView view{b->seq1.data(), b->off1.data(), b->seq2.data(), b->off2.data(), b->n_pairs,
                           b->pair_method.empty() ? nullptr : b->pair_method.data()};
        std::string err;
        pa_config cfg = b->cfg;
        if ((st = pa::make_plan(cfg, view, forced.data(), &sat, pd->plan, err)) != PA_OK) return fail(ctx, st, err);
        pd->subset_host = sat;
        if ((st = upload_plan(ctx, *pd)) != PA_OK) return st;
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
    e = pa::launch_init(ia, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "init launch");
    return launch_plan(ctx, b, *b->recompute, false);
}

// The wavefront kernel's bounded waits set flags[n_pairs] when they give up.
pa_status check_error_slot(pa_ctx *ctx, pa_batch *b)
{
    uint32_t err = 0;
    PA_HIP(hipMemcpy(&err, static_cast<uint32_t *>(b->flags.ptr) + b->n_pairs, 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (err) return fail(ctx, PA_ERR_DEVICE, "wavefront kernel: a boundary wait exceeded its limit");
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

pa_status pa_batch_create(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                          const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs,
                          const uint8_t *pair_method, pa_batch **out)
{
    if (!ctx || !cfg || !out || !off1 || !off2 || (n_pairs && (!seq1 || !seq2)))
        return fail(ctx, PA_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    auto t0 = std::chrono::steady_clock::now();
    auto b = std::make_unique<pa_batch>();
    b->cfg = *cfg;
    if (cfg->score_kind == PA_SCORE_MATRIX && cfg->matrix && cfg->matrix_symbols && cfg->matrix_dim > 0) {
        b->matrix.assign(cfg->matrix, cfg->matrix + size_t(cfg->matrix_dim) * cfg->matrix_dim);
        b->symbols.assign(cfg->matrix_symbols, cfg->matrix_symbols + cfg->matrix_dim);
        b->cfg.matrix = b->matrix.data();
        b->cfg.matrix_symbols = b->symbols.data();
    }
    b->n_pairs = n_pairs;
    b->off1.assign(off1, off1 + n_pairs + 1);
    b->off2.assign(off2, off2 + n_pairs + 1);
    b->seq1.assign(seq1 ? seq1 : nullptr, seq1 ? seq1 + off1[n_pairs] : nullptr);
    b->seq2.assign(seq2 ? seq2 : nullptr, seq2 ? seq2 + off2[n_pairs] : nullptr);
    if (pair_method) b->pair_method.assign(pair_method, pair_method + n_pairs);
    if (b->seq1.empty()) b->seq1.push_back(0);
    if (b->seq2.empty()) b->seq2.push_back(0);

    pa::BatchView view{b->seq1.data(), b->off1.data(), b->seq2.data(), b->off2.data(), n_pairs,
                       pair_method ? b->pair_method.data() : nullptr};
    std::string err;
    pa_status st = pa::make_plan(b->cfg, view, nullptr, nullptr, b->main.plan, err);
    if (st != PA_OK) return fail(ctx, st, err);
    const pa::Plan &p = b->main.plan;
    // Only int16 pairs whose bound does not exclude saturation need the flag read-back.
    b->can_saturate = p.n_speculative > 0;

    if ((st = upload_plan(ctx, b->main)) != PA_OK) return st;
    if ((st = upload(ctx, b->pairs, p.pairs)) != PA_OK) return st;
    PA_HIP(hipMalloc(&b->scores.ptr, std::max<size_t>(n_pairs * 4, 16)), "hipMalloc");
    PA_HIP(hipMalloc(&b->flags.ptr, (n_pairs + 1) * 4), "hipMalloc");   // + error slot
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");

    b->stats.n_pairs = n_pairs;
    b->stats.cells = p.cells;
    b->stats.padded_cells = p.padded_cells;
    b->stats.n_int16 = p.n16;
    b->stats.n_int32 = p.n32;
    b->stats.n_launches = p.launches.size() + 1;
    b->stats.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = b.release();
    return PA_OK;
}

pa_status pa_batch_run(pa_ctx *ctx, pa_batch *b)
{
    if (!ctx || !b) return PA_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PA_ERR_DEVICE, "hipSetDevice");
    pa_status st = run_pass(ctx, b, false);
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
    b->main_ms_accum = 0;
    b->main_ms_count = 0;
    PA_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    PA_HIP(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
    for (int i = 0; i < iters; ++i) {
        pa_status st = run_pass(ctx, b, ms_main != nullptr);
        if (st != PA_OK) return st;
    }
    PA_HIP(hipEventRecord(ctx->ev[1], ctx->stream), "hipEventRecord");
    PA_HIP(hipEventSynchronize(ctx->ev[1]), "hipEventSynchronize");
    float ms = 0;
    PA_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
    if (ms_total) *ms_total = ms;
    if (ms_main) *ms_main = b->main_ms_count ? b->main_ms_accum / double(b->main_ms_count) : 0.0;
    return check_error_slot(ctx, b);
}

pa_status pa_batch_destroy(pa_ctx *ctx, pa_batch *b)
{
    if (!b) return PA_ERR_INVALID_ARGUMENT;
    if (ctx) (void)hipSetDevice(ctx->device);
    delete b;
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

} // extern "C"
