// pa_plan.hpp -- host-side planning for the striped DP kernels (plain C++17, no HIP).
//
// The planner turns a batch of pairs plus a pa_config into
//   * per-pair metadata (sizes after orientation, method, lane width),
//   * tasks: one task = one group of G lanes sweeping one pair (int32) or two pairs
//     packed into the 16-bit halves of each 32-bit lane (int16),
//   * transposed, pre-encoded residue streams (one 32-bit word per DP row and per DP column of
//     a task, the two halves holding the two pairs), so every load in the kernel is one dword,
//   * launch groups: contiguous task ranges that share one kernel variant.
//
// The residue streams are encoded on the device (pa_encode.hip) from the caller's bytes, which are
// uploaded once per batch as one buffer: the seq1 bytes [0, end1) followed by the seq2 bytes.
//
// Layout of one task in HBM (offsets are in 32-bit words):
//   rowcodes[row_off + s*G*R + t*R + r]  residue of DP row i = s*G*R + t*R + r + 1 (stripe s,
//                                         lane t, register r), padded to S*G*R rows
//   colcodes[col_off + j]                residue of DP column j+1, padded to col_words(N) words
//   bnd[2*(bnd_off + j) + {0,1}]          H and vertical-gap state below the current stripe at
//                                         column j+1 (multi-stripe tasks only)
#pragma once

#include <cstdint>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "pa_device.h"

namespace pa {

inline constexpr uint32_t kNone = 0xFFFFFFFFu;

enum ScoreCode : uint8_t { kScoreDna = 0, kScoreMatrix = 1, kScorePairs = 2, kScoreProfile = 3 };

// Residue encodings.  Unitary scoring over at most four symbols uses a byte-permute lookup:
// score = perm(tbl_hi, tbl_lo, rowcode ^ colcode), where the xor of the two codes yields the
// selector bytes of the 16-bit score for "equal" (entry 0) or "different" (entries 1..3).
//   int16 half : row (2c)|(2c<<8),           column (2c)|((2c+1)<<8)
//   int32 word : row 2c|2c<<8|c<<16|c<<24,   column 2c|(2c+1)<<8|(8+c)<<16|(8+c)<<24
//                (selector bytes 8..11 replicate the sign of the selected 16-bit entry)
// Padding codes make every pad cell score -1 (selector bytes >= 13 give 0xFF).
inline constexpr uint32_t kDnaRowPad16 = 0x1010u, kDnaColPad16 = 0x2020u;
inline constexpr uint32_t kDnaRowPad32 = 0x10101010u, kDnaColPad32 = 0x20202020u;

inline uint32_t dna_row16(uint32_t c) { return (2 * c) | ((2 * c) << 8); }
inline uint32_t dna_col16(uint32_t c) { return (2 * c) | ((2 * c + 1) << 8); }
inline uint32_t dna_row32(uint32_t c) { return (2 * c) | ((2 * c) << 8) | (c << 16) | (c << 24); }
inline uint32_t dna_col32(uint32_t c) { return (2 * c) | ((2 * c + 1) << 8) | ((8 + c) << 16) | ((8 + c) << 24); }

// Matrix scoring: an int16 table [alphabet][stride] (stride even); the symbol of rank K (the
// alphabet size without the pad) is the pad symbol and scores -1 against everything.  In LDS
// the table is replicated 32 times so that lane l of every half-wave reads only bank l % 32:
// entry e lives at byte ((e >> 1) * 32 + l % 32) * 4 + (e & 1) * 2.  With e = a*stride + b and
// an even stride that address splits into a row part, a column part and a lane part:
//   row code = a * (stride/2) * 128,  column code = (b >> 1) * 128 + (b & 1) * 2,
//   lane part = (l % 32) * 4 (added by the kernel when it loads the row codes).
// int16 tasks keep the two halves' codes in the two 16-bit halves of one word.
inline constexpr uint32_t kMaxMatrixSymbols = 31;   // plus the pad: 32 x 16 x 128 B = 64 KiB of LDS
inline uint32_t lds_row_offset(uint32_t a, uint32_t stride) { return a * (stride / 2) * 128; }
inline uint32_t lds_col_offset(uint32_t b) { return (b >> 1) * 128 + (b & 1) * 2; }
inline uint32_t lds_table_bytes(uint32_t alphabet, uint32_t stride) { return alphabet * (stride / 2) * 128; }

// Pair-row table (int16 tasks, alphabets of at most kPairTableMaxAlphabet symbols with the pad):
// one dword per (a1, a2, b) holding s(a1, b) in the low and s(a2, b) in the high half, so one
// LDS read gives the scores of two consecutive rows (a1, a2) of one pair against column residue
// b.  Stored once (42.6 KB for 22 symbols); an int16 task's row code is one word per two rows,
// the 16-bit byte offset (a1*K + a2)*K*4 of each half, its column code b*4 per half.
inline constexpr uint32_t kPairTableMaxAlphabet = 22;
inline uint32_t pair_row_offset(uint32_t a1, uint32_t a2, uint32_t K) { return (a1 * K + a2) * K * 4; }
inline uint32_t pair_col_offset(uint32_t b) { return b * 4; }
inline uint32_t pair_table_bytes(uint32_t K) { return K * K * K * 4; }

// Query profile (int16 tasks of one query against many targets; striped kernel): the column code
// of target residue rank b is the byte offset b * rows * 2 of its profile column, and the lane
// reads its R consecutive rows from there.  Rows per lane 8 and 16 are left out: their lanes'
// reads would fall on the same LDS banks.
inline constexpr uint32_t kProfileMaxBytes = 65536;
inline constexpr uint32_t kProfileRowsMask = 0x7F & ~0x44u;   // R = 4, 6, 10, 12, 14
inline uint32_t profile_col_offset(uint32_t b, uint32_t rows) { return b * rows * 2; }

// Allocator whose value-less construct leaves trivially constructible elements uninitialised:
// the planner's large per-task arrays are sized first and then filled (and paged in) by
// parallel passes, instead of being zeroed on one thread.
template <typename T>
struct UninitAllocator : std::allocator<T> {
    template <typename U> struct rebind { using other = UninitAllocator<U>; };
    UninitAllocator() noexcept = default;
    template <typename U> UninitAllocator(const UninitAllocator<U> &) noexcept {}
    template <typename U> void construct(U *p) noexcept { ::new (static_cast<void *>(p)) U; }
    template <typename U, typename... A> void construct(U *p, A &&...a)
    {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <typename T> using FillVector = std::vector<T, UninitAllocator<T>>;

struct Task {
    uint32_t row_off;
    uint32_t col_off;
    uint32_t bnd_off;   // kNone if S == 1
    uint32_t N;         // DP columns swept (max over the halves)
    uint32_t S;         // stripes of G*R rows
    uint32_t pair[2];   // output index per half, kNone if the half is unused
    uint32_t m[2];      // DP rows of each half (>= 1)
    uint32_t n[2];      // DP columns of each half (>= 1)
    uint32_t bias;      // global int16 three-way-maximum variants (max3_variant): the bias of half 0
                        // in the low and of half 1 in the high 16 bits (int16_bias); else 0
};
static_assert(sizeof(Task) == 48, "Task layout is shared with the device");

// Where the device encoder reads a task's residues: per half, the byte offsets of the row and the
// column sequence in the batch's residue buffer (seq2 bytes start at BatchView::end1()).
struct EncodeTask {
    uint64_t row_src[2];
    uint64_t col_src[2];
};
static_assert(sizeof(EncodeTask) == 32, "EncodeTask layout is shared with the device");

struct VariantKey {
    uint8_t width;   // 16 (two pairs per lane, saturating int16) or 32 (one pair per lane, int32)
    uint8_t G;       // lanes per group
    uint8_t R;       // rows per lane
    uint8_t score;   // ScoreCode
    uint8_t local;   // method_local
    uint8_t track;   // int16 global: keep the running maximum to detect saturation
    uint8_t rowmax;  // global with trailing gaps free in both dimensions: per-row maxima
    uint8_t export_dp;   // also write the final DP row and column of every pair (int32, R = 8)
    bool operator<(const VariantKey &o) const;
    bool operator==(const VariantKey &o) const;
};
static_assert(sizeof(VariantKey) == 8, "VariantKey is compared as one 64-bit word by the planner");

// Substitution-score shift of a kernel variant, in units of -ge (pa_cells.hpp): the global
// method drifts along the rows (1) and, unless it keeps per-row maxima, along the columns as
// well (2); the local method drifts along the columns only (1), except in the DP export, which
// keeps exact values (0).
inline constexpr int kScoreShifts = 3;
inline int score_shift(const VariantKey &key)
{
    if (key.local) return key.export_dp ? 0 : 1;   // column drift (not in the DP export)
    return (key.track || key.rowmax || key.export_dp) ? 1 : 2;
}

// Rows per lane of the DP-export variants (int32 lanes, striped kernel only).
inline constexpr uint8_t kExportRows = 8;

struct Launch {
    VariantKey key;
    uint32_t task_begin, task_end;
    uint64_t padded_cells;   // lane-cells swept, both int16 halves counted
};

// A fused launch: its leader launch and its wave items [items_begin, items_end) in
// Plan::fused_items, in decreasing order of modelled time.
struct FusedLaunch {
    uint32_t leader;
    uint32_t items_begin, items_end;
    double share;   // its members' modelled work over the pass's (local cells weighed 10/8)
};

// Wave item of a fused launch: up to 64 / 16 consecutive tasks of one member launch, as
// (R - 4) / 2 in bits 29-31, the task count less one in bits 27-28 and the first task's index in
// the plan's task array in bits 0-26.
inline constexpr uint32_t kFusedMaxTasks = 1u << 27;
inline uint32_t fused_item(uint32_t R, uint32_t count, uint32_t first)
{
    return ((R - 4) / 2) << 29 | (count - 1) << 27 | first;
}

// Launches that a fused launch can take: the int16 striped variants without per-row maxima or the
// DP export, of the local method with the byte-permute scores or the pair-row table, and of the
// global method with the byte-permute scores (fused in launch form 4 only).
inline bool fused_launchable(const VariantKey &k)
{
    if (k.width != 16 || k.G != 16 || k.track || k.rowmax || k.export_dp) return false;
    return k.local ? (k.score == kScoreDna || k.score == kScorePairs) : k.score == kScoreDna;
}

// The rows-per-lane range of the fused kernel that takes a fused_launchable launch: every R for
// the local method; for the global method R <= 12 and R >= 14 apart, two kernels of 254 and 256
// VGPRs, two waves per SIMD like its sweeps from R = 10 up alone (R = 16 holds 234 alone; one
// kernel holding R = 12 to 16 needs 274, every R 331 -- one wave per SIMD).
struct RowsRange { int lo, hi; };
inline RowsRange fused_rows(const VariantKey &k)
{
    if (k.local) return {4, 16};
    return k.R <= 12 ? RowsRange{4, 12} : RowsRange{14, 16};
}

struct PairMeta {
    uint32_t m, n;        // after orientation: m rows (seq1 side), n columns
    uint8_t local;
    uint8_t swapped;      // rows/columns exchanged relative to the caller's seq1/seq2
    uint8_t width;        // 0 (empty: closed form), 16 or 32
    uint8_t pad_;
};

// Planner and launch parameters resolved from pa_tuning (zero fields take these defaults).
struct Tuning {
    double cell_cost = 13.0;     // ~VALU instructions per row of a lane per step
    double step_cost = 24.0;     // per-step work outside the cells (lane exchange, conveyors, loop)
    uint32_t rows_mask = 0x7F;   // bit i: R = 4 + 2i allowed
    uint32_t wavefront_min_rows = 3072;
    // a pair beyond wavefront_min_rows runs in the wavefront kernel only if its modelled striped
    // sweep exceeds this multiple of the batch's striped work per resident lane group (0: always)
    double wavefront_tail = 0.75;
    uint32_t streams = 4;
    bool eager = false;
    bool cost_order = false;
    // Launch forms of the striped launches.  0 (default): one block per 16 tasks, except that the
    // int16 local launches that differ only in R run as one fused launch (resident blocks, one
    // work queue; fused_launchable); 1: every striped launch runs resident blocks taking tasks
    // from a work queue of its own; 2: as 1, with the fused_launchable launches fused as in 0;
    // 3: one block per 16 tasks, nothing fused; 4: as 0, with the global method's int16
    // byte-permute launches also fused, in two ranges of R (fused_rows).
    uint32_t persistent = 0;
};
Tuning resolve_tuning(const pa_tuning *t);

struct Plan {
    Tuning tuning;
    // scoring
    int32_t go = 0, ge = 0;
    int32_t fc = 0, fr = 0, lc = 0, lr = 0;   // after orientation
    uint8_t score = kScoreDna;
    // Every table exists in kScoreShifts versions, version k with the scores shifted by -k*ge
    // (score_shift: the drift the kernel variant adds per diagonal step).
    uint32_t tbl_lo[kScoreShifts] = {}, tbl_hi[kScoreShifts] = {};   // DNA byte-permute table
    uint32_t stride = 0, alphabet = 0;        // matrix table geometry (alphabet includes the pad)
    std::vector<int16_t> table[kScoreShifts];      // matrix: alphabet*stride entries
    std::vector<uint32_t> ptable[kScoreShifts];    // pair-row tables (empty: alphabet too large)
    // Query profile (one query against many targets, int16): entry [b * profile_rows + i] is the
    // score of query residue i against residue rank b, -1 for pad rows and the pad rank; two
    // entries per word.  Empty: not used.
    std::vector<uint32_t> profile[kScoreShifts];
    uint32_t profile_rows = 0;                // query rows covered (padded for every allowed R)
    uint32_t profile_ranks = 0;               // residue ranks including the pad
    int32_t smax = 0, smin = 0;               // extreme substitution scores
    int32_t lbias = 0;                        // int16_local_bias of this configuration

    std::vector<PairMeta> pairs;
    FillVector<Task> tasks;
    std::vector<Launch> launches;
    uint64_t bnd_entries = 0;                 // (H,E) entries of the boundary buffer
    uint64_t cells = 0, padded_cells = 0;
    uint64_t n16 = 0, n32 = 0;
    uint64_t n_speculative = 0;               // int16 pairs whose bound does not exclude saturation
    uint64_t n_profiled = 0;                  // int16 pairs scored through the query profile
    uint64_t n_empty = 0;                     // planned pairs with an empty sequence (no task)
    FillVector<EncodeTask> enc;               // per task: residue sources for the device encoder
    // DP export: per task, the index in the output of H(0, n) (the task's last column, m+1
    // entries) and of H(m, 0) (its last row, n+1 entries).  The output holds the caller's
    // dp_column values (seq1 side, pair p at off1[p]-off1[0]+p) followed by the dp_row values.
    FillVector<uint64_t> dp_base;             // 2 per task
    uint64_t dp_entries = 0;
    uint64_t n_rowcodes = 0, n_colcodes = 0;  // words of the encoded residue streams
    int16_t code[256] = {};                   // residue -> DNA code or matrix rank (-1: absent)
    uint32_t pad_rank = 0;                    // matrix rank of the pad symbol
    std::vector<uint32_t> launch_order;       // launches in start order (order_launches)
    std::vector<uint8_t> launch_stream;       // stream of each launch (indexed like launches)
    // Fused launches (Tuning::persistent 0 or 2): one persistent launch runs the tasks of several
    // launches whose keys differ only in R.  It takes its leader's place in launch_order, stream,
    // work queue and arguments; the other members are not launched on their own.
    std::vector<FusedLaunch> fused;
    std::vector<int32_t> fused_of;            // per launch: index into fused, or -1
    std::vector<uint32_t> fused_items;        // wave items of all fused launches (fused_item)
};

struct BatchView {
    const uint8_t *seq1;
    const uint64_t *off1;
    const uint8_t *seq2;
    const uint64_t *off2;
    uint64_t n_pairs;
    const uint8_t *pair_method;  // nullable
    // One query against many targets: every pair's first sequence is seq1[off1[0], off1[1]) (off1
    // holds two entries).
    bool one_query = false;
    // Residue codes of an earlier plan of the same batch (Plan::code): the alphabet is taken from
    // them and the residue bytes are not read (seq1 only for a one-query profile, seq2 not at all).
    const int16_t *known_codes = nullptr;

    const uint8_t *s1(uint64_t p) const { return seq1 + off1[one_query ? 0 : p]; }
    uint64_t src1(uint64_t p) const { return off1[one_query ? 0 : p]; }   // offset of s1(p) in seq1
    uint64_t len1(uint64_t p) const { return one_query ? off1[1] - off1[0] : off1[p + 1] - off1[p]; }
    uint64_t end1() const { return off1[one_query ? 1 : n_pairs]; }   // end of the first sequences
};

// Widths chosen per pair before packing: 16 or 32 (0 = let the policy decide).
// `only` restricts planning to a subset of pairs (nullptr = all).
// `track_all` keeps the running maximum of H in every global variant (labels for the learned
// range predictor).
// `pair_cost` (nullable): a predicted cost per pair (the learned dispatcher's cost output); with
// it the tasks of a launch are ordered by it instead of by the geometric model.
pa_status make_plan(const pa_config &cfg, const BatchView &batch, const uint8_t *forced_width,
                    const std::vector<uint32_t> *only, Plan &plan, std::string &err, bool track_all = false,
                    bool export_dp = false, const float *pair_cost = nullptr);

// Modelled sweep cost of one task (arbitrary units): stripes x steps x per-step work.
double task_cost(const VariantKey &key, const Task &t, const Tuning &tuning);

// Largest substitution score of a configuration (int16 range analysis and model features).
int32_t max_substitution(const pa_config &cfg);

// Number of colcodes / bnd words reserved per task: the kernels prefetch up to four
// conveyor chunks of G words past the last column.
inline uint32_t col_words(uint32_t N, uint32_t G) { return ((N + G - 1) / G) * G + 4 * G; }

// Re-centred int16 wavefront tasks (global, no per-row maxima): the kernel keeps each wave's
// values relative to an offset moved every chunk, so int16 needs only the spread of the cells in
// flight to fit: adjacent cells differ by at most kRecentreStep(...) and the cells of a wave lie
// within kRecentreReach steps of each other.  Their HBM boundary rows hold exact int32 values
// (two entries per column).
inline constexpr uint32_t kRecentreReach = 64 * 16 + 3 * 64;
inline bool recentred(const VariantKey &key)
{
    return key.G == 64 && key.width == 16 && !key.local && !key.track && !key.rowmax;
}

// Conservative int16 range analysis of one pair (rows m, columns n).
struct RangeBound {
    int64_t lower;   // every real H and gap state is >= lower
    int64_t upper;   // every real H is <= upper
};
RangeBound range_bound(const Plan &plan, bool local, uint32_t m, uint32_t n);

// int16 variants with the column drift take the maximum of three values in one instruction, the
// packed fp16 three-input maximum (v_pk_maximum3_f16), whose order on the 16-bit patterns
// 0x0000..0x7C00 (zero, the positive numbers, +infinity) is their integer order.  Patterns
// 0x7C01..0x7FFF are NaNs (the instruction returns a NaN) and 0x8000..0xFFFF are negative (in
// reverse order; -1023..-1 are NaNs too).  The kernels therefore hold the values of these
// variants biased so that every operand of a three-way maximum is in [0, kInt16Top] -- or,
// for values derived from the low boundary words of columns past a task's end, at most -1024,
// below every operand in range -- and a value above kInt16Top is an int16 overflow: it turns
// the pair's running maximum into a NaN pattern, which the saturation test flags.
inline constexpr int32_t kInt16Top = 31744;
inline bool max3_variant(const VariantKey &key)
{
    return key.width == 16 && !key.track && !key.rowmax && !key.export_dp && !recentred(key);
}
// Bias of the local method's max3 variants, the same for every pair of a plan: every operand of
// a three-way maximum is >= min(smin, go + ge) in exact terms (H >= 0; the diagonal term adds a
// substitution score, the pad's -1 included; gap states are >= H + go + ge), plus a margin of
// |ge| + 1 that also keeps the running maximum's start (the bias in the frame of step -1) >= 0.
int32_t int16_local_bias(int32_t smin, int32_t go, int32_t ge);
// Bias of one pair of a max3 variant: the local bias, or for the global method one that lifts
// the pair's lower bound (less one substitution score for the diagonal term) to >= 0.  The
// variant fits the pair when upper + bias <= kInt16Top (int16_fits).
int32_t int16_bias(const Plan &plan, bool local, uint32_t m, uint32_t n);
inline bool int16_fits(const RangeBound &rb, int32_t bias) { return rb.upper + bias <= kInt16Top; }
// Largest local score the max3 int16 lanes compute exactly for this configuration (the learned
// dispatcher's limit): kInt16Top less the local bias and the largest drift, |ge| * 81.
int32_t int16_score_limit(const pa_config &cfg);

// Pairs with more rows than this go to the wavefront kernel (G = 64, a workgroup per task).
inline constexpr uint32_t kWavefrontMinRows = 3072;

// Choice of (G, R) for a pair with m rows: G = 64, R = 16 for the wavefront kernel, else G = 16
// and the allowed rows per lane with the lowest modelled sweep cost (stripes x steps x per-step
// work; the step count is the same for every R).
void choose_shape(uint32_t m, const Tuning &tuning, bool wavefront, uint8_t &G, uint8_t &R);

// Modelled cost of an m x n pair swept by one lane group of the striped kernel (G = 16, the rows
// per lane choose_shape picks): S * (n + G - 1) * (cell_cost * R + step_cost).
double striped_cost(uint32_t m, uint32_t n, const Tuning &tuning);

// Pairs with more rows than this (two rounds of the wavefront kernel's four 1,024-row stripes)
// go to the wavefront kernel whatever the rest of the batch: its waves are then busy for all but
// the last round's remainder, and (int32) it sweeps faster than a lane group of the striped kernel.
inline constexpr uint32_t kWavefrontAlwaysRows = 8 * 1024;

// Lane groups the GPU keeps resident at once in the model: 1,024 SIMDs x 3 waves x 4 groups.
inline constexpr double kResidentGroups = 1024.0 * 3 * 4;

// LDS bytes a launch of variant `key` needs for its substitution table.
uint32_t lds_bytes(const Plan &plan, const VariantKey &key);

// How the device encoder forms the residue codes of a launch of variant `key`: the row and column
// code of every byte value (for the pair-row table the row table holds ranks) and the pad codes.
struct CodeTables {
    uint32_t row[256];
    uint32_t col[256];
    uint32_t row_pad, col_pad;
};
void code_tables(const Plan &plan, const VariantKey &key, CodeTables &out);

// Groups the plan's fused_launchable launches that differ only in R into fused launches
// (Plan::fused, fused_of, fused_items), unless Tuning::persistent is 1 or 3.  Wave items come in
// decreasing order of modelled time, or with pair_cost (nullable) of predicted cost.
void fuse_launches(Plan &plan, const float *pair_cost = nullptr);

// Order of the launches of a plan and the stream of each: wavefront launches first (longest
// task first), then striped launches by the time each would take on its own (a fused launch
// once, in its leader's place).  Wavefront launches take streams of their own, striped launches
// the others round robin.
void order_launches(Plan &plan);

} // namespace pa
