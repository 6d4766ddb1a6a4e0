/*
 * pa_device.h -- C ABI of the MI355X (gfx950) batched affine-gap score-only aligner.
 *
 * The reference (rrahn/pairwise_aligner) is a header-only C++20 template library with no ABI:
 * an aligner is configured by a compile-time rule chain and called with compute().  This
 * header is the plain-C boundary that sits exactly where the reference's compute() calls
 * its DP (interface_one_to_one_bulk.hpp:73-100 -> dp_algorithm_template_standard.hpp:43-100).
 * The header-only C++ host layer (include/pairwise_aligner/...) and the Python mirror
 * (pairwise_aligner_amd) both translate the reference's configuration into pa_config and
 * call the functions below.  No C++ or torch types cross this boundary.
 *
 * Conventions: seq1 indexes DP rows, seq2 DP columns (as in the reference).  Sequences are
 * passed as one concatenated byte buffer per side with n_pairs+1 uint64 offsets.  Scores are
 * returned as int32 with the reference's scalar (int32) semantics; a local alignment with an
 * empty sequence returns INT32_MIN (local_affine_scalar_test.cpp:100-105).
 */
#ifndef PA_DEVICE_H
#define PA_DEVICE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PA_ABI_VERSION 7

typedef enum pa_status {
    PA_OK = 0,
    PA_ERR_INVALID_ARGUMENT = 1,  /* null pointer, bad enum, inconsistent offsets */
    PA_ERR_INVALID_SYMBOL = 2,    /* a residue outside the substitution-matrix alphabet */
    PA_ERR_UNSUPPORTED = 3,       /* gap model or alphabet the device path does not handle */
    PA_ERR_DEVICE = 4,            /* HIP runtime error (message in pa_ctx_last_error) */
    PA_ERR_OUT_OF_MEMORY = 5,
    PA_ERR_RANGE = 6              /* a score could leave the int32 range */
} pa_status;

/* method: cfg::method_global / cfg::method_local (method_global.hpp:35-53, method_local.hpp:41-49) */
enum { PA_METHOD_GLOBAL = 0, PA_METHOD_LOCAL = 1 };
/* score model: cfg::score_model_unitary* / cfg::score_model_matrix* (score_model_unitary.hpp:52-59,
 * score_model_matrix.hpp:52-64) */
enum { PA_SCORE_UNITARY = 0, PA_SCORE_MATRIX = 1 };
/* lane policy: which integer width the DP runs in.  AUTO = int16 wherever a bound derived from
 * lengths and scores proves it cannot saturate, int32 elsewhere; LEARNED = int16 wherever the
 * learned range predictor says the scores fit; INT16 = always try int16 first.  Every int16
 * result that reaches the int16 limit is recomputed at int32 (PA_PAIR_RECOMPUTED). */
enum { PA_LANES_AUTO = 0, PA_LANES_LEARNED = 1, PA_LANES_INT16 = 16, PA_LANES_INT32 = 32 };

/* Planner and launch parameters: the autotuner's search space (pa_tune).  No counterpart in the
 * reference, whose tile and lane shapes are compile-time constants (dp_matrix_lane_width.hpp:25-55).
 * A zero field takes the default given in brackets. */
typedef struct pa_tuning {
    float cell_cost;              /* shape model: relative cost of one row of one lane per step [13] */
    float step_cost;              /* shape model: relative fixed cost of one lane step [24] */
    uint32_t rows_mask;           /* rows per lane R in {4,6,...,16} allowed: bit R/2-2 [all] */
    uint32_t wavefront_min_rows;  /* pairs with more DP rows use the wavefront kernel [3072] */
    uint32_t streams;             /* concurrent launch streams, 1..4 [4] */
    uint32_t eager;               /* 1: launch kernels one by one; 0: replay a captured hipGraph [0] */
    uint32_t cost_order;          /* PA_LANES_LEARNED: 1 orders the waves of a launch (runs of tasks formed in
                                     the geometric order) by the network's predicted cost instead of the
                                     geometric model's [0] */
    uint32_t persistent;          /* launch form of the striped launches.  0: one block per 16 tasks, except
                                     that the int16 local launches that differ only in their rows per lane
                                     run as one launch of resident blocks taking tasks from one work queue;
                                     1: every striped launch runs as many blocks as the GPU keeps resident,
                                     whose waves take tasks from a work queue of its own; 2: as 1, with the
                                     int16 local launches fused as in 0; 3: one block per 16 tasks,
                                     nothing fused; 4: as 0, with the global method's int16 launches
                                     of the unitary (DNA) scores also fused, in two launches by rows
                                     per lane [0] */
    float wavefront_tail;         /* a pair beyond wavefront_min_rows (and within 8,192 rows) runs in the wavefront
                                     kernel only if its modelled striped sweep exceeds this multiple of the batch's
                                     striped work per resident lane group (where the batch has less other work than
                                     that); < 0: every such pair [0.75] */
} pa_tuning;

typedef struct pa_config {
    int32_t method;             /* PA_METHOD_* */
    int32_t first_column_free;  /* cfg::leading_end_gap::first_column == end_gap::free */
    int32_t first_row_free;     /* cfg::leading_end_gap::first_row */
    int32_t last_column_free;   /* cfg::trailing_end_gap::last_column */
    int32_t last_row_free;      /* cfg::trailing_end_gap::last_row */
    int32_t gap_open;           /* cfg::gap_model_affine(open, extension): a gap of k costs open + k*extension */
    int32_t gap_extension;
    int32_t score_kind;         /* PA_SCORE_* */
    int32_t match;              /* unitary */
    int32_t mismatch;
    int32_t matrix_dim;         /* matrix: number of symbols (<= 31) */
    const int32_t *matrix;      /* matrix: dim*dim scores, [rank of seq1 symbol][rank of seq2 symbol] */
    const uint8_t *matrix_symbols; /* matrix: the dim symbols; rank = position (alphabet_rank_map_scalar.hpp:39-51) */
    int32_t lane_policy;        /* PA_LANES_* */
    const struct pa_model *model; /* PA_LANES_LEARNED: the trained range predictor (else ignored) */
    const pa_tuning *tuning;    /* planner/launch parameters; NULL = defaults */
} pa_config;

/* per-pair result flags */
#define PA_PAIR_INT16 0x01      /* final score computed in packed int16 lanes */
#define PA_PAIR_INT32 0x02      /* final score computed in int32 lanes */
#define PA_PAIR_RECOMPUTED 0x04 /* an int16 pass reached the int16 limit; recomputed at int32 */
#define PA_PAIR_EMPTY 0x08      /* an empty sequence: closed-form score */

typedef struct pa_ctx pa_ctx;
typedef struct pa_batch pa_batch;
typedef struct pa_model pa_model;

typedef struct pa_batch_stats {
    uint64_t n_pairs;
    uint64_t cells;             /* sum of m*n over the batch (interior cells) */
    uint64_t padded_cells;      /* cells the kernels actually sweep, padding included */
    uint64_t n_int16;           /* pairs planned in int16 lanes */
    uint64_t n_int32;           /* pairs planned in int32 lanes */
    uint64_t n_recomputed;      /* pairs recomputed at int32 after the last run */
    uint64_t n_launches;        /* kernel launches per run */
    double plan_ms;             /* host planning time */
    uint64_t n_profiled;        /* int16 pairs scored through a query profile (one-to-many) */
} pa_batch_stats;

/* Context: one HIP device and one HIP stream.  Not thread-safe; use one per thread. */
pa_status pa_ctx_create(pa_ctx **out, int device);
pa_status pa_ctx_destroy(pa_ctx *ctx);
const char *pa_ctx_last_error(const pa_ctx *ctx);
int pa_abi_version(void);

/* One-shot: upload, plan, run, download.  Replaces aligner.compute(seqs1, seqs2)
 * (interface_one_to_one_bulk.hpp:47-100 and, for one pair, interface_one_to_one_single.hpp:40-67).
 * pair_method (nullable) overrides cfg->method per pair.  flags (nullable) receives PA_PAIR_*. */
pa_status pa_score_batch(pa_ctx *ctx, const pa_config *cfg,
                         const uint8_t *seq1, const uint64_t *off1,
                         const uint8_t *seq2, const uint64_t *off2,
                         uint64_t n_pairs, const uint8_t *pair_method,
                         int32_t *scores, uint8_t *flags);

/* The same plus the final DP column and row of every pair: the scores of the reference's
 * result.dp_column() / dp_row() (result/aligner_result.hpp:41-59, filled by
 * dp_algorithm_template_standard.hpp:43-100; the vectors the global tracker reads,
 * tracker_global_scalar.hpp:36-61).  dp_column receives H(i, len2), i = 0..len1, of pair p at
 * (off1[p] - off1[0]) + p; dp_row receives H(len1, j), j = 0..len2, at (off2[p] - off2[0]) + p.
 * Runs every pair in int32 lanes of the striped kernel. */
pa_status pa_score_batch_dp(pa_ctx *ctx, const pa_config *cfg,
                            const uint8_t *seq1, const uint64_t *off1,
                            const uint8_t *seq2, const uint64_t *off2,
                            uint64_t n_pairs, const uint8_t *pair_method,
                            int32_t *scores, uint8_t *flags, int32_t *dp_column, int32_t *dp_row);

/* pa_score_batch_dp with whole cells: the reference's DP vectors hold affine cells (score, gap
 * state) (affine/affine_cell.hpp:25-73), updated by compute_cell (affine_dp_algorithm.hpp:139-156).
 * dp_column_gap and dp_row_gap (both or neither; the layout of dp_column / dp_row) receive the gap
 * halves: W(i, len2 + 1) = max(W(i, len2) + ge, H(i, len2) + go + ge), the horizontal gap state
 * right of the last column, for i >= 1, and V(len1 + 1, j), the vertical one below the last row,
 * for j >= 1; entry 0 of each vector keeps its initial go + ge (the sweep never updates it,
 * dp_algorithm_template_base.hpp:94-128).  Without a sequence, a vector keeps its initial cells
 * (H, H + go + ge) (affine_initialisation_strategy.hpp:38-51). */
pa_status pa_score_batch_dp_cells(pa_ctx *ctx, const pa_config *cfg,
                                  const uint8_t *seq1, const uint64_t *off1,
                                  const uint8_t *seq2, const uint64_t *off2,
                                  uint64_t n_pairs, const uint8_t *pair_method,
                                  int32_t *scores, uint8_t *flags, int32_t *dp_column, int32_t *dp_row,
                                  int32_t *dp_column_gap, int32_t *dp_row_gap);

/* One-shot scoring that allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Arguments, results and errors as pa_score_batch, except:
 *   - every device buffer of the call is a slice of one static array per device of
 *     PA_NOALLOC_SCRATCH_BYTES (64 MiB), which the library holds from the first pa_ctx_create on
 *     that device until the process ends; a batch whose buffers do not fit returns
 *     PA_ERR_OUT_OF_MEMORY (the message says how much was needed);
 *   - PA_LANES_LEARNED returns PA_ERR_UNSUPPORTED; launches are eager (no graph is instantiated);
 *   - calls on one device share its array, which keeps the last call's contents: calls that
 *     overlapped on one device would race, so the library serialises them with a per-device lock
 *     (other processes hold their own copy).  The call is synchronous; it is not capturable into
 *     a HIP graph. */
#define PA_NOALLOC_SCRATCH_BYTES (64u << 20)
pa_status pa_score_batch_noalloc(pa_ctx *ctx, const pa_config *cfg,
                                 const uint8_t *seq1, const uint64_t *off1,
                                 const uint8_t *seq2, const uint64_t *off2,
                                 uint64_t n_pairs, const uint8_t *pair_method,
                                 int32_t *scores, uint8_t *flags);

/* Device-resident batch: the same work split into plan/upload, repeated runs, and download,
 * so a caller can time runs with inputs already in HBM.  The input buffers are read only during
 * the call: the batch keeps the residues on the device and no host copy of them. */
pa_status pa_batch_create(pa_ctx *ctx, const pa_config *cfg,
                          const uint8_t *seq1, const uint64_t *off1,
                          const uint8_t *seq2, const uint64_t *off2,
                          uint64_t n_pairs, const uint8_t *pair_method, pa_batch **out);
/* One query against many targets.  Replaces aligner.compute(seq1, bulk2)
 * (interface_one_to_many_bulk.hpp:47-92) and its query profile (score_model_matrix_simd_1xN.hpp:63-183,
 * dp_matrix_lane_profile.hpp:53-80): scores[p] is the score of (query, seq2[p]), exactly as
 * pa_score_batch with the query repeated as every seq1.  When the query is the row side of every
 * pair (every configuration except a free last row with a penalised last column), fits the striped
 * kernel and its profile fits 64 KiB (about 1,400 residues for a 21-letter matrix, 6,500 for DNA),
 * the int16 tasks read their scores from a query profile in LDS instead of the residue tables. */
pa_status pa_score_one_to_many(pa_ctx *ctx, const pa_config *cfg,
                               const uint8_t *query, uint64_t query_len,
                               const uint8_t *seq2, const uint64_t *off2,
                               uint64_t n_targets, const uint8_t *pair_method,
                               int32_t *scores, uint8_t *flags);
pa_status pa_batch_create_one_to_many(pa_ctx *ctx, const pa_config *cfg,
                                      const uint8_t *query, uint64_t query_len,
                                      const uint8_t *seq2, const uint64_t *off2,
                                      uint64_t n_targets, const uint8_t *pair_method, pa_batch **out);
/* Enqueue every kernel of one pass on the context stream and, when int16 results reached
 * the limit, the int32 recomputation.  Returns when the pass is complete. */
pa_status pa_batch_run(pa_ctx *ctx, pa_batch *batch);
pa_status pa_batch_results(pa_ctx *ctx, pa_batch *batch, int32_t *scores, uint8_t *flags);
pa_status pa_batch_stats_get(const pa_batch *batch, pa_batch_stats *out);
/* Runs the batch `iters` times.  *ms_total is the HIP-event time of all passes on the context
 * stream; *ms_main is the HIP-event time of the main DP kernel launches alone (averaged per
 * pass), used for the roofline. */
pa_status pa_batch_time(pa_ctx *ctx, pa_batch *batch, int iters, double *ms_total, double *ms_main);
pa_status pa_batch_destroy(pa_ctx *ctx, pa_batch *batch);

/* Autotuner.  No counterpart in the reference.  Plans and times candidate tunings and writes
 * the fastest to *out, in two stages of coordinate descent: the shape-model fields (step_cost,
 * rows_mask) on a sample of the batch (every k-th pair, about sample_pairs pairs; 0 = 5% of the
 * batch, at least 2,000), then the launch-level fields (wavefront_min_rows, streams, eager,
 * persistent, wavefront_tail) and the stage-1 result against the defaults on the whole batch;
 * a final pick that differs from cfg's tuning is kept only if it is also faster in two more
 * alternating timings against it.  `iters` passes are timed per candidate (0 = 3). */
pa_status pa_tune(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                  const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, const uint8_t *pair_method,
                  uint64_t sample_pairs, int iters, pa_tuning *out);

/* ---------------------------------------------------------------------------------------------
 * Learned lane dispatch.  No counterpart in the reference (its lane width is a compile-time
 * choice of the score-model rule, score_model_unitary_simd.hpp:51-137).  A 3-layer FP32 MLP
 * (PA_FEATURES -> 64 -> 64 -> 2) maps per-pair features to
 *   y0 = max over the DP of H / (max substitution score * min(m, n))   (score range)
 *   y1 = log2(modelled sweep cost of the pair) / 32                     (cell cost)
 * Feature extraction, forward, backward and the Adam update run on the GPU.
 * ------------------------------------------------------------------------------------------- */
#define PA_FEATURES 8
#define PA_MODEL_OUTPUTS 2
#define PA_MODEL_HIDDEN 64

/* Creates a model with seeded random weights (He initialisation, zero biases). */
pa_status pa_model_create(pa_ctx *ctx, uint64_t seed, pa_model **out);
pa_status pa_model_destroy(pa_model *model);
int pa_model_param_count(void);
/* Copies the flat parameter vector [W1 (F x H), b1, W2 (H x H), b2, W3 (H x 2), b3] in or out. */
pa_status pa_model_set_params(pa_ctx *ctx, pa_model *model, const float *params);
pa_status pa_model_get_params(pa_ctx *ctx, const pa_model *model, float *params);
/* Per-pair features (n_pairs x PA_FEATURES, row-major) of a batch, computed on the GPU:
 * log lengths, length ratio, hashed 3-mer composition distance, k-mer containment and the
 * identity it implies, substitution score scale. */
pa_status pa_features(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                      const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, float *features);
/* Forward pass: y (n x 2). */
pa_status pa_model_forward(pa_ctx *ctx, const pa_model *model, const float *x, uint64_t n, float *y);
/* Mean-squared-error loss and its gradient with respect to the parameters (flat, as above);
 * y weights (n x 2, nullable) scale each output's error. */
pa_status pa_model_loss_grad(pa_ctx *ctx, const pa_model *model, const float *x, const float *target,
                             const float *weight, uint64_t n, float *loss, float *grad);
/* The relative safety margin of the lane choice: a pair runs in int16 lanes when its predicted
 * maximum of H times (1 + margin), plus 256, stays below the largest score the int16 lanes hold
 * exactly (31,744 less the configuration's value bias and the largest drift of its frames; else
 * int32).  A new model has 0.10; a trainer sets it from the network's error on held-out pairs. */
pa_status pa_model_set_margin(pa_model *model, float margin);
/* One Adam step on a minibatch; returns the loss before the step. */
pa_status pa_model_train_step(pa_ctx *ctx, pa_model *model, const float *x, const float *target,
                              const float *weight, uint64_t n, float learning_rate, float *loss);
/* Per-pair labels produced by the DP itself: the maximum of H over the matrix (int32 pass with
 * the running maximum kept) and the modelled sweep cost of the pair's task in the plan. */
pa_status pa_dp_labels(pa_ctx *ctx, const pa_config *cfg, const uint8_t *seq1, const uint64_t *off1,
                       const uint8_t *seq2, const uint64_t *off2, uint64_t n_pairs, int32_t *max_h, float *cost);

#ifdef __cplusplus
}
#endif

#endif /* PA_DEVICE_H */
