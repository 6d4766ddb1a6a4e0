/*
 * gotoh_oracle.c -- CPU oracle for batched affine-gap (Gotoh) score-only alignment.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the device library is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product path (pairwise_aligner_amd, include/) never links,
 * loads or calls anything built from this directory.
 *
 * It restates, in plain C with 64-bit accumulators, the scalar algorithm of
 * rrahn/pairwise_aligner:
 *
 *   cell recurrence          include/pairwise_aligner/affine/affine_dp_algorithm.hpp:139-156
 *   boundary initialisation  include/pairwise_aligner/affine/affine_initialisation_strategy.hpp:38-51
 *   end-gap flags            include/pairwise_aligner/configuration/end_gap_policy.hpp:20-36
 *   local method             include/pairwise_aligner/configuration/method_local.hpp:41-49
 *   local zero clamp         include/pairwise_aligner/matrix/dp_matrix_local.hpp:39-47
 *   global extraction        include/pairwise_aligner/tracker/tracker_global_scalar.hpp:36-61
 *   local extraction         include/pairwise_aligner/tracker/tracker_local_scalar.hpp:38-49
 *   unitary score            include/pairwise_aligner/score_model/score_model_unitary.hpp:52-59
 *   matrix score             include/pairwise_aligner/score_model/score_model_matrix.hpp:52-64
 *   rank map                 include/pairwise_aligner/alphabet_conversion/alphabet_rank_map_scalar.hpp:39-63
 *
 * Semantics (seq1 indexes rows i = 1..m, seq2 indexes columns j = 1..n):
 *   H(0,0) = 0
 *   H(i,0) = first_column free ? 0 : go + ge*i            (i >= 1)
 *   H(0,j) = first_row    free ? 0 : go + ge*j            (j >= 1)
 *   V(i,j) = max(V(i-1,j) + ge, H(i-1,j) + go + ge)       vertical gap state, V(1,j) = H(0,j)+go+ge
 *   W(i,j) = max(W(i,j-1) + ge, H(i,j-1) + go + ge)       horizontal gap state, W(i,1) = H(i,0)+go+ge
 *   D(i,j) = H(i-1,j-1) + s(a_i, b_j)       local: D = max(D, 0)
 *   H(i,j) = max(D, V, W)
 *   global score = H(m,n), maxed with every H(m,j) (j = 0..n) if last_row is free and
 *                  with every H(i,n) (i = 0..m) if last_column is free;
 *   local score  = max of H over interior cells (i,j >= 1), starting from INT32_MIN,
 *                  so an empty sequence yields INT32_MIN (local_affine_scalar_test.cpp:100-105).
 *   local forces all four end gaps free (method_local.hpp:41-49).
 *
 * The final DP column and row hold affine cells (score, gap state) (affine/affine_cell.hpp:25-73),
 * initialised as (H, H + go + ge) (affine_initialisation_strategy.hpp:38-51) and updated by
 * compute_cell (affine_dp_algorithm.hpp:139-156) whenever the cell is on the path of the sweep
 * (dp_algorithm_template_base.hpp:82-129, which never touches entry 0 of either vector):
 *   column gap state, i >= 1:  W(i, n+1) = max(W(i,n) + ge, H(i,n) + go + ge);   i = 0: go + ge
 *   row gap state,    j >= 1:  V(m+1, j) = max(V(m,j) + ge, H(m,j) + go + ge);   j = 0: go + ge
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define PAO_GLOBAL 0
#define PAO_LOCAL 1
#define PAO_UNITARY 0
#define PAO_MATRIX 1

typedef struct pao_config {
    int32_t mode;               /* PAO_GLOBAL / PAO_LOCAL */
    int32_t first_column_free;  /* leading gaps of seq1 free: H(i,0) = 0 */
    int32_t first_row_free;     /* leading gaps of seq2 free: H(0,j) = 0 */
    int32_t last_column_free;   /* max over the last column H(i,n) */
    int32_t last_row_free;      /* max over the last row H(m,j) */
    int32_t gap_open;
    int32_t gap_extension;
    int32_t score_kind;         /* PAO_UNITARY / PAO_MATRIX */
    int32_t match;
    int32_t mismatch;
    int32_t dim;                /* matrix: number of symbols */
    const int32_t *matrix;      /* matrix: dim*dim, [rank(seq1 symbol)][rank(seq2 symbol)] */
    const uint8_t *symbols;     /* matrix: dim symbols, rank = position in this list */
} pao_config;

static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* Rank map of alphabet_rank_map_scalar.hpp:39-51: 256 entries, -1 for symbols not in the list. */
static void build_rank_map(const pao_config *cfg, int32_t map[256])
{
    for (int c = 0; c < 256; ++c) map[c] = -1;
    for (int r = 0; r < cfg->dim; ++r) map[cfg->symbols[r]] = r;
}

/* Returns 0 on success, -1 if a symbol is outside the matrix alphabet.  last_row (n+1 entries,
 * nullable) receives H(m,j), j = 0..n, and last_col (m+1 entries, nullable) H(i,n), i = 0..m: the
 * scores of the reference's final DP row and column (aligner_result.hpp:41-59; the vectors the
 * global tracker reads, tracker_global_scalar.hpp:36-61).  last_row_gap / last_col_gap (nullable,
 * same sizes) receive the gap-state halves of the same cells (see the header). */
int pao_score_dp(const uint8_t *s1, int64_t m, const uint8_t *s2, int64_t n,
                 const pao_config *cfg, int64_t *out_score, int64_t *last_row, int64_t *last_col,
                 int64_t *last_row_gap, int64_t *last_col_gap)
{
    const int local = cfg->mode == PAO_LOCAL;
    const int fc = local || cfg->first_column_free;
    const int fr = local || cfg->first_row_free;
    const int lc = local || cfg->last_column_free;
    const int lr = local || cfg->last_row_free;
    const int64_t go = cfg->gap_open, ge = cfg->gap_extension;

    int32_t map[256];
    int32_t *r1 = NULL, *r2 = NULL;
    if (cfg->score_kind == PAO_MATRIX) {
        build_rank_map(cfg, map);
        r1 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
        r2 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        for (int64_t i = 0; i < m; ++i) { r1[i] = map[s1[i]]; if (r1[i] < 0) { free(r1); free(r2); return -1; } }
        for (int64_t j = 0; j < n; ++j) { r2[j] = map[s2[j]]; if (r2[j] < 0) { free(r1); free(r2); return -1; } }
    }

    /* Row buffers over columns j = 0..n: H of the previous row and the vertical gap state. */
    int64_t *Hrow = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t *Vrow = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    Hrow[0] = 0;
    for (int64_t j = 1; j <= n; ++j) Hrow[j] = fr ? 0 : go + ge * j;
    for (int64_t j = 0; j <= n; ++j) Vrow[j] = Hrow[j] + go + ge;

    int64_t best_local = INT32_MIN;
    int64_t best_col = Hrow[n];          /* max over H(i,n), i = 0..m */
    if (last_col) last_col[0] = Hrow[n];
    if (last_col_gap) last_col_gap[0] = go + ge;

    for (int64_t i = 1; i <= m; ++i) {
        int64_t hcol0 = fc ? 0 : go + ge * i;  /* H(i,0) */
        int64_t diag = Hrow[0];                 /* H(i-1,0) */
        int64_t w = hcol0 + go + ge;            /* W(i,1) */
        Hrow[0] = hcol0;
        for (int64_t j = 1; j <= n; ++j) {
            int64_t s;
            if (cfg->score_kind == PAO_MATRIX)
                s = cfg->matrix[(int64_t)r1[i - 1] * cfg->dim + r2[j - 1]];
            else
                s = (s1[i - 1] == s2[j - 1]) ? cfg->match : cfg->mismatch;
            int64_t d = diag + s;
            if (local) d = max64(d, 0);
            int64_t v = Vrow[j];                /* V(i,j) */
            int64_t h = max64(max64(d, v), w);
            diag = Hrow[j];
            Hrow[j] = h;
            if (local) best_local = max64(best_local, h);
            int64_t open = h + go + ge;
            Vrow[j] = max64(v + ge, open);      /* V(i+1,j) */
            w = max64(w + ge, open);            /* W(i,j+1) */
        }
        best_col = max64(best_col, Hrow[n]);
        if (last_col) last_col[i] = Hrow[n];
        if (last_col_gap) last_col_gap[i] = w;   /* W(i,n+1) */
    }
    if (last_row) for (int64_t j = 0; j <= n; ++j) last_row[j] = Hrow[j];
    /* V(m+1,j); entry 0 keeps its initial go + ge (Vrow[0] is never updated) */
    if (last_row_gap) for (int64_t j = 0; j <= n; ++j) last_row_gap[j] = Vrow[j];

    int64_t score;
    if (local) {
        score = best_local;
    } else {
        score = Hrow[n];
        if (lr) for (int64_t j = 0; j <= n; ++j) score = max64(score, Hrow[j]);
        if (lc) score = max64(score, best_col);
    }
    free(Hrow); free(Vrow); free(r1); free(r2);
    *out_score = score;
    return 0;
}

int pao_score(const uint8_t *s1, int64_t m, const uint8_t *s2, int64_t n,
              const pao_config *cfg, int64_t *out_score)
{
    return pao_score_dp(s1, m, s2, n, cfg, out_score, NULL, NULL, NULL, NULL);
}

typedef struct batch_job {
    const uint8_t *seq1; const uint64_t *off1;
    const uint8_t *seq2; const uint64_t *off2;
    const pao_config *cfg;
    const uint8_t *pair_mode;   /* optional per-pair PAO_GLOBAL/PAO_LOCAL override */
    int64_t *scores;
    int64_t *rows, *cols;       /* optional final DP rows / columns (pao_dp_batch) */
    int64_t *row_gaps, *col_gaps;   /* optional gap-state halves of the same cells */
    int64_t begin, end;
    int status;
} batch_job;

static void *batch_worker(void *arg)
{
    batch_job *job = (batch_job *)arg;
    pao_config local_cfg = *job->cfg;
    for (int64_t p = job->begin; p < job->end; ++p) {
        if (job->pair_mode) local_cfg.mode = job->pair_mode[p];
        const uint8_t *a = job->seq1 + job->off1[p];
        const uint8_t *b = job->seq2 + job->off2[p];
        int64_t m = (int64_t)(job->off1[p + 1] - job->off1[p]);
        int64_t n = (int64_t)(job->off2[p + 1] - job->off2[p]);
        int64_t *row = job->rows ? job->rows + (job->off2[p] - job->off2[0]) + p : NULL;
        int64_t *col = job->cols ? job->cols + (job->off1[p] - job->off1[0]) + p : NULL;
        int64_t *rg = job->row_gaps ? job->row_gaps + (job->off2[p] - job->off2[0]) + p : NULL;
        int64_t *cg = job->col_gaps ? job->col_gaps + (job->off1[p] - job->off1[0]) + p : NULL;
        if (pao_score_dp(a, m, b, n, &local_cfg, &job->scores[p], row, col, rg, cg) != 0) job->status = -1;
    }
    return NULL;
}

/* Scores n_pairs pairs given as concatenated sequences with n_pairs+1 offsets and, when rows /
 * cols are given, their final DP rows (pair p at (off2[p]-off2[0]) + p, n_p+1 entries) and columns
 * (pair p at (off1[p]-off1[0]) + p, m_p+1 entries); row_gaps / col_gaps (nullable) receive the
 * gap-state halves in the same layout.  Splits the batch over n_threads POSIX threads.  Returns 0
 * or -1 (invalid symbol). */
int pao_dp_batch(const uint8_t *seq1, const uint64_t *off1,
                 const uint8_t *seq2, const uint64_t *off2,
                 int64_t n_pairs, const pao_config *cfg, const uint8_t *pair_mode,
                 int64_t *scores, int64_t *rows, int64_t *cols, int64_t *row_gaps, int64_t *col_gaps,
                 int n_threads)
{
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    batch_job jobs[256];
    pthread_t tids[256];
    int64_t chunk = (n_pairs + n_threads - 1) / n_threads;
    int started = 0;
    for (int t = 0; t < n_threads; ++t) {
        int64_t b = t * chunk, e = b + chunk;
        if (b >= n_pairs) break;
        if (e > n_pairs) e = n_pairs;
        jobs[t] = (batch_job){seq1, off1, seq2, off2, cfg, pair_mode, scores, rows, cols, row_gaps, col_gaps, b, e, 0};
        if (n_threads == 1) { batch_worker(&jobs[t]); started = 1; break; }
        pthread_create(&tids[t], NULL, batch_worker, &jobs[t]);
        ++started;
    }
    int status = 0;
    for (int t = 0; t < started; ++t) {
        if (n_threads > 1) pthread_join(tids[t], NULL);
        if (jobs[t].status) status = -1;
    }
    return status;
}

int pao_score_batch(const uint8_t *seq1, const uint64_t *off1,
                    const uint8_t *seq2, const uint64_t *off2,
                    int64_t n_pairs, const pao_config *cfg, const uint8_t *pair_mode,
                    int64_t *scores, int n_threads)
{
    return pao_dp_batch(seq1, off1, seq2, off2, n_pairs, cfg, pair_mode, scores, NULL, NULL, NULL, NULL, n_threads);
}
