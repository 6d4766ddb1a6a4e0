// pa_kernels.hpp -- launch interface of the gfx950 kernels (host side of pa_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_plan.hpp"

namespace pa {

// Device-side flag bit (internal): an int16 pass reached the int16 limit.
inline constexpr uint32_t kFlagSaturated = 0x100u;

struct KernelArgs {
    const Task *tasks;
    uint32_t ntasks;
    const uint32_t *rowcodes;
    const uint32_t *colcodes;
    uint32_t *bnd;
    int32_t *scores;
    uint32_t *flags;            // per pair; flags[error_slot] is set when a bounded wait gives up
    uint32_t error_slot;
    uint32_t self_init;         // the DP kernels write their pairs' initial scores and flags
    int32_t *maxh;              // nullable: per-pair maximum of H over the matrix (LOCAL or TRACK variants)
    const int16_t *table;
    uint32_t table_entries;
    int32_t go, ge;
    int32_t lbias;              // int16 max3 variants of the local method: the bias of their values
    uint32_t tbl_lo, tbl_hi;
    int32_t fc, fr, lc, lr;
    int32_t *dp_out;            // DP export variants: the final rows and columns (scores), then
                                // the gap-state halves of the same cells, dp_gap entries further on
    const uint64_t *dp_base;    // 2 per task of this launch (pa_plan.hpp)
    uint64_t dp_gap;
    uint32_t *queue;            // nullable: the launch's work queue (next task, finished blocks), zero
                                // between passes; striped kernel only
    // fused launch (pa_plan.hpp, FusedLaunch): its wave items; tasks index the plan's whole task
    // array (tasks points at its first task)
    const uint32_t *items;
    uint32_t nitems;
    float share;                // fused launch: the share of the GPU's resident blocks it takes
};

struct InitArgs {
    const PairMeta *pairs;
    const uint32_t *subset;   // nullable: only these pair indices
    uint64_t n_pairs;         // pairs, or entries of subset
    int32_t *scores;
    uint32_t *flags;
    int32_t go, ge;
    int32_t fc, fr, lc, lr;
    uint32_t flag_or;         // extra flag bits written with the initial value
};

// Residue-stream encoder (pa_encode.hip): writes the row and column codes of tasks
// [task_begin, task_end) of one launch from the batch's residue bytes.
struct EncodeArgs {
    const Task *tasks;          // the plan's tasks
    const EncodeTask *enc;      // per task: residue sources
    uint32_t task_begin, task_end;
    const uint8_t *residues;    // seq1 bytes, then seq2 bytes
    const CodeTables *codes;    // this launch's code tables (device copy)
    uint32_t *rowcodes, *colcodes;
    uint32_t stripe_rows;       // G * R
    uint32_t G;
    uint32_t width;             // 16 or 32
    uint32_t score;             // ScoreCode
    uint32_t pair_K;            // pair-row table: residue ranks with the pad
};

hipError_t launch_stripe(const VariantKey &key, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st);
// A fused launch of the variants that differ from `key` only in R (pa_plan.hpp, fused_launchable):
// a.items / a.nitems its wave items, a.queue its work queue.
hipError_t launch_fused(const VariantKey &key, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st);
hipError_t launch_encode(const EncodeArgs &a, hipStream_t st);
hipError_t launch_init(const InitArgs &a, hipStream_t st);

} // namespace pa
