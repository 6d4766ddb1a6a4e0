// pa_dp_kernels.hpp -- gfx950 DP kernels of the batched affine-gap (Gotoh) score-only aligner and
// their per-variant launch dispatch.  Instantiated per (lane width, rows per lane) in
// pa_kernels_inst.hip, which the build compiles once per pair so that the variants compile in
// parallel; pa_kernels.hip routes a launch to its instantiation.
//
// Both DP kernels are systolic sweeps over the DP matrix.  A lane owns R consecutive DP rows
// of the current stripe and keeps their H and horizontal-gap state in registers.  The lanes of
// a group move along the DP columns skewed by one step per lane: at step k lane t computes
// column c = k - t for its R rows, receiving from lane t-1 by DPP the H and vertical-gap state
// of the row above its block and the column residue.  Lane 0 receives the stripe's top
// boundary instead: the initial DP row for the first stripe, otherwise the bottom row of the
// stripe above.  Column residues and boundary values reach lane 0 through "conveyors": one
// chunk of consecutive words loaded coalesced, one word per lane, then shifted towards lane 0
// by one lane per step.
//
//   stripe_kernel     short pairs.  A group is one DPP row of G = 16 lanes (row_shr:1) and
//                     sweeps one task: one pair in int32 lanes or two pairs in the two 16-bit
//                     halves of every lane (packed int16, v_pk_add_i16 clamp / v_pk_max_i16).
//                     Stripes of 16*R rows run one after the other; the boundary row between
//                     them goes through HBM.
//   wavefront_kernel  long pairs.  A workgroup of 4 waves sweeps one task; every wave is one
//                     64-lane group (wave_shr:1) owning a stripe of 64*R rows, and the waves
//                     take the stripes round robin, so 4 stripes are in flight at once, each
//                     wave a fixed distance behind the one above.  The boundary row between
//                     consecutive waves passes through an LDS ring buffer with progress
//                     counters; from the last wave back to the first it goes through HBM.
//
// Boundary (affine_initialisation_strategy.hpp:38-51): H(i,0) = first_column free ? 0 : go+ge*i,
// H(0,j) likewise with first_row; the gap state entering the matrix is H + go + ge.
// Score extraction (tracker_global_scalar.hpp:36-61, tracker_local_scalar.hpp:38-49):
//   global  H(m,n) is captured by the lane owning row m at the step of column n; with a free
//           last column every lane contributes max over its real rows at column n; with a
//           free last row the per-row maxima are kept and the owner of row m contributes its
//           row's maximum.  Contributions meet in an int32 atomicMax.
//   local   running maximum over all cells.  Pad rows and columns score -1 against
//           everything, so with gap scores <= 0 no pad cell exceeds the maximum of the real
//           cells.
// Saturation: int16 lanes saturate; a running maximum equal to 32767 flags the pair for an
// int32 recomputation.  Lower bounds are proven on the host before int16 is chosen.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "pa_cells.hpp"
#include "pa_kernels.hpp"
#include "pa_plan.hpp"

namespace pa {

// Top-boundary value for columns past the end of a pair: low enough that no real cell is
// raised by it, high enough that adding gap penalties cannot wrap in int32 lanes
// (int16 lanes read it as two halves of -16384).
constexpr uint32_t kLowWord = 0xC000C000u;

// Group reduction of the running maximum over 16 lanes (one DPP row).
template <int W>
__device__ __forceinline__ typename Lane<W>::V row_reduce_max(typename Lane<W>::V v)
{
    using L = Lane<W>;
    uint32_t m = L::bits(v);
    m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor8>(m, m))));
    m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor4>(m, m))));
    m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor2>(m, m))));
    m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor1>(m, m))));
    return L::from(m);
}

// Initial score of a pair with m, n >= 1 (tracker_global_scalar.hpp:36-61): INT32_MIN, raised by
// the free-end candidates that no DP cell visits, H(m, 0) with a free last row and H(0, n) with a
// free last column (affine_initialisation_strategy.hpp:38-51).
__device__ __forceinline__ int32_t initial_score(int m, int n, int go, int ge, int fc, int fr, int lc, int lr)
{
    int64_t s = INT32_MIN;
    if (lr) { const int64_t h = fc ? 0 : go + int64_t(ge) * m; s = h > s ? h : s; }
    if (lc) { const int64_t h = fr ? 0 : go + int64_t(ge) * n; s = h > s ? h : s; }
    return int32_t(s);
}

// With a.self_init: the initial score and flags of a task's pairs, written by one lane before
// any other lane of the wave (or, with the barrier that follows, of the workgroup) raises the
// score with an atomic or flags a saturation.
template <int W, bool LOCAL>
__device__ __forceinline__ void init_task_pairs(const KernelArgs &a, const Task &tk, int fc, int fr, int lc, int lr)
{
#pragma unroll
    for (int q = 0; q < Lane<W>::NH; ++q) {
        if (tk.pair[q] == kNone) continue;
        a.flags[tk.pair[q]] = W == 16 ? 0x01u : 0x02u;
        // the local maximum starts at INT32_MIN (the wavefront kernel's waves raise it atomically)
        a.scores[tk.pair[q]] =
            LOCAL ? INT32_MIN : initial_score(int(tk.m[q]), int(tk.n[q]), a.go, a.ge, fc, fr, lc, lr);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}

// One task of the striped kernel, swept by the lane group of G lanes that holds it (t = the lane's
// index in its group): one stripe after the other, the boundary row between stripes through HBM.
// The launch's substitution table is in LDS (lds_raw) for the score kinds that read one.
// EXPORT: also write the task's final DP column and row (int32 lanes; pa_plan.hpp, dp_base).
// A task of one stripe (m <= G*R; the group's choice, uniform over its 16 lanes) takes a sweep
// without boundary rows through HBM, whose lane 0 keeps its top boundary (the initial DP row) in
// two registers instead of conveyors.  Tasks are ordered by decreasing work, so the one- and
// multi-stripe tasks of a launch share at most one wave.
template <int W, int G, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX, bool EXPORT>
__device__ __forceinline__ void stripe_task(const KernelArgs &a, const uint32_t gid, const char *lds_raw)
{
    using L = Lane<W>;
    using V = typename L::V;
    const int t = int(threadIdx.x % G);
    const uint32_t lane_bank = (threadIdx.x & 31u) * 4u;   // this lane's copy of the LDS table
    const Task tk = a.tasks[gid];
    const int go = a.go, ge = a.ge;
    // method_local.hpp:41-49: a local alignment has every end gap free
    const int fc = LOCAL ? 1 : a.fc, fr = LOCAL ? 1 : a.fr, lc = LOCAL ? 0 : a.lc;
    if (a.self_init && t == 0) init_task_pairs<W, LOCAL>(a, tk, fc, fr, lc, LOCAL ? 0 : a.lr);
    constexpr bool DRIFT = !LOCAL;           // row drift (pa_cells.hpp) for the global method
    constexpr bool ACC = ROWMAX || TRACK;    // per-row maxima
    // column drift: on top of the row drift for the global method, alone for the local method
    // (whose zero clamp and running maximum then work in the step's frame)
    constexpr bool CDRIFT = !ACC && !EXPORT;
    // int16 lanes with the column drift take three-way maxima in one instruction and hold
    // their values biased into its range (pa_plan.hpp, kInt16Top): the local method by one
    // bias for the launch, the global method by one per half, from the task
    constexpr bool MAX3 = W == 16 && CDRIFT;
    const int lbias = MAX3 && LOCAL ? a.lbias : 0;
    const int b0 = MAX3 && !LOCAL ? int(tk.bias & 0xFFFFu) : lbias;
    const int b1 = MAX3 && !LOCAL ? int(tk.bias >> 16) : lbias;
    const V vbias = L::pair(b0, b1);

    const int dge = DRIFT ? ge : 0;
    const int cg = CDRIFT ? -ge : 0;         // column drift per step
    const V vgo = L::splat(go), vgogo = L::splat(go + ge), vge = L::splat(ge);
    // shifts the last row's drift back when H, V leave the lane (and, with the column drift,
    // forward by one step: the next lane uses them one step later)
    const V vout = L::splat((DRIFT ? ge * R : 0) + (CDRIFT ? -ge : 0));
    const V vrebase = L::splat(ge * G);      // end of a chunk: the column drift drops by G|ge|
    uint32_t tbl_lo = a.tbl_lo;
    const uint32_t tbl_hi = a.tbl_hi;
    // the permute's second table word must be a VGPR (one scalar operand per instruction): keep
    // it in one for the whole kernel instead of moving it from its scalar register every step
    if constexpr (SCORE == kScoreDna) asm volatile("" : "+v"(tbl_lo));
    const int N = int(tk.N);
    constexpr int ROWS = G * R;
    static_assert(!EXPORT || W == 32, "the DP export runs in int32 lanes");
    int32_t *dp_col = nullptr, *dp_row = nullptr;   // H(0, n) and H(m, 0) of this task's vectors
    if constexpr (EXPORT) {
        dp_col = a.dp_out + a.dp_base[2 * gid];
        dp_row = a.dp_out + a.dp_base[2 * gid + 1];
        if (t == 0) {
            dp_col[0] = h0_value(int(tk.n[0]), fr, go, ge);
            dp_row[0] = h0_value(int(tk.m[0]), fc, go, ge);
            // entry 0 of each vector keeps its initial gap state, go + ge: the sweep never
            // updates it (dp_algorithm_template_base.hpp:94-128 reads entries from 1 on)
            dp_col[a.dp_gap] = go + ge;
            dp_row[a.dp_gap] = go + ge;
        }
    }

    // exact; for the local method with the column drift, 0 in the frame of step -1 (it moves to
    // the frame of each step as the step begins and back by G steps at each chunk end)
    V runmax = LOCAL ? L::splat((CDRIFT ? -cg : 0) + lbias) : L::lowest();
    V satmax = L::lowest();                         // drifted (int16 saturation test)
    const V vcg = L::splat(cg);

    // one stripe of the task; ONE: the task has no other
    auto stripe = [&]<bool ONE>(uint32_t s) {
        const int row0 = int(s) * ROWS + t * R;   // DP row above this lane's block
        Rows<W, R, ACC> st;
        // lane t starts at step t: its initial column (j = 0) counts as computed at step t-1
        init_rows<W, R, SCORE, ACC, DRIFT, CDRIFT>(st, a.rowcodes + tk.row_off, row0, fc, go, ge, lane_bank,
                                                   cg * (t - 1), vbias);
        V Hdiag = L::add(L::splat(h0_value(row0, fc, go, ge) + cg * (t - 1)), vbias);
        Capture<W, R> cap;
        // the step at which this lane's next capture event comes (events are rare: one wave
        // test per step against it)
        int kev = INT32_MAX;
        if constexpr (!LOCAL) {
            cap.init(tk, row0, lc);
            kev = cap.ev == INT32_MAX ? INT32_MAX : cap.ev + t;
        }

        const bool has_bnd_in = !ONE && s > 0;
        const bool has_bnd_out = !ONE && s + 1 < tk.S;
        const uint32_t *colp = a.colcodes + tk.col_off;
        uint32_t *bndp = has_bnd_in || has_bnd_out ? a.bnd + 2ull * tk.bnd_off : nullptr;

        // Column-code conveyor: chunk j holds the residues of columns jG+1 .. jG+G, one per
        // lane, so at the start of step k lane 0 of `conv` holds the residue of column k+1;
        // `nextB` is the chunk after it.  The residue of step k+1 is thus known one step early,
        // and its substitution scores are read while the cells of step k are computed.
        uint32_t Bin = colp[t];   // step 0: lane 0 computes column 0
        uint32_t conv = colp[1 + t];
        uint32_t nextB = colp[G + 1 + t];
        // Top-boundary conveyor: chunk j holds H and the vertical gap state above this stripe
        // for columns jG .. jG+G-1: the initial row H(0, c+1) for the first stripe, the stored
        // bottom row of the stripe above otherwise.  Columns past N (computed, never read, by
        // the unmasked steps) get a low value that no real cell can be raised by.
        // Lane t of a chunk is consumed at the chunk's step t: column drift cg*t.
        auto top_chunk = [&](int col, uint32_t &h, uint32_t &e) {
            if (has_bnd_in) {
                if (col < N) {
                    h = bndp[2 * col];
                    e = bndp[2 * col + 1];
                    if constexpr (CDRIFT) {
                        h = L::bits(L::add(L::from(h), L::splat(cg * t)));
                        e = L::bits(L::add(L::from(e), L::splat(cg * t)));
                    }
                } else {
                    h = e = kLowWord;
                }
            } else {
                // H(0, col+1) exact, the vertical gap state below it (drifted by |ge| for row 0)
                const int hv = h0_value(col + 1, fr, go, ge);
                h = L::bits(L::add(L::from(L::top_word(hv + cg * t)), vbias));
                e = L::bits(L::add(L::from(L::top_word(hv + go + ge - dge + cg * t)), vbias));
            }
        };
        uint32_t convH = 0, convE = 0, nextH = 0, nextE = 0;
        // ONE: lane 0's top boundary at the current step, H(0, c+1) and the vertical gap state
        // below it; it changes by top_step per step (the free first row's column drift, or the
        // penalised first row's extension without the column drift) and by the chunk rebase
        uint32_t topH = 0, topE = 0;
        const int top_step = CDRIFT ? (fr ? cg : 0) : (fr ? 0 : ge);
        const V vtop_step = L::splat(top_step);
        if constexpr (ONE) {
            top_chunk(0, topH, topE);
        } else {
            top_chunk(t, convH, convE);
            top_chunk(G + t, nextH, nextE);
        }
        // outputs of this lane from the previous step (read by lane t+1)
        uint32_t Hout = 0, Eout = 0;
        // bottom-row outputs of lane G-1 over one chunk of G steps, shifted down one lane per
        // step: at the end of the chunk lane j holds the output of the chunk's step j
        uint32_t qH = 0, qE = 0;

        ScoreReads<W, R, SCORE> rd;
        issue_scores<W, R, SCORE>(st, Bin, lds_raw, rd);
        V sc[R];
        finish_scores<W, R, SCORE>(st, Bin, tbl_lo, tbl_hi, rd, sc);

        // One step.  MASKED: lanes before their first column (k - t < 0; the first chunk,
        // where lane t waits t steps) leave their state alone, and so do lanes past column N-1
        // in the ROWMAX/TRACK/EXPORT variants, whose per-row maxima and exported rows must not
        // see pad columns.  Elsewhere lanes past column N-1 sweep pad columns, as in every
        // unmasked step: pad cells never exceed the real ones and never feed them, and a lane's
        // state never falls behind the column drift (a state left alone while the chunk ends
        // drop the drift would sink by G|ge| per chunk, for the max3 variants into patterns
        // their three-way maximum does not order).
        // TOPSTEP (ONE only): lane 0's top boundary changes from step to step (top_step != 0)
        auto step = [&]<bool MASKED, bool TOPSTEP>(int k) {
            const int c = k - t;   // DP column (0-based) of this lane at this step
            // the conveyor moves on first, so that the old register is free for lane 0's value
            const uint32_t shifted = dpp_any<kRowShl1>(conv);
            const uint32_t Bnext = dpp<kRowShr1>(conv, Bin);
            conv = shifted;
            issue_scores<W, R, SCORE>(st, Bnext, lds_raw, rd);
            uint32_t Hin, Ein;
            if constexpr (ONE) {
                // lanes 1.. take lane t-1's outputs, lane 0 keeps its top boundary
                topH = dpp<kRowShr1>(topH, Hout);
                topE = dpp<kRowShr1>(topE, Eout);
                Hin = topH;
                Ein = topE;
                if constexpr (TOPSTEP) {
                    topH = L::bits(L::add(L::from(topH), vtop_step));
                    topE = L::bits(L::add(L::from(topE), vtop_step));
                }
            } else {
                const uint32_t nextH_ = dpp_any<kRowShl1>(convH), nextE_ = dpp_any<kRowShl1>(convE);
                Hin = dpp<kRowShr1>(convH, Hout);
                Ein = dpp<kRowShr1>(convE, Eout);
                convH = nextH_;
                convE = nextE_;
            }
            if constexpr (LOCAL && CDRIFT) runmax = L::add(runmax, vcg);   // into this step's frame
            if (!MASKED || (ROWMAX || TRACK || EXPORT ? uint32_t(c) < uint32_t(N) : c >= 0)) {
                V Eup = L::from(Ein);
                // EXPORT: the vertical gap state below the last row, from the lane that holds it
                V vlast{};
                const int er = EXPORT ? int(tk.m[0]) - 1 - row0 : -1;
                update_cells<W, R, SCORE, LOCAL, ACC, DRIFT, CDRIFT, MAX3, EXPORT>(
                    st, Eup, Hdiag, runmax, sc, Bin, tbl_lo, tbl_hi, vgo, vgogo, vge,
                    L::splat(cg * (k % G) + lbias), &vlast, er);
                Hdiag = L::from(Hin);
                Hout = L::bits(DRIFT || CDRIFT ? L::add(st.H[R - 1], vout) : st.H[R - 1]);
                Eout = L::bits(DRIFT || CDRIFT ? L::add(Eup, vout) : Eup);
                if constexpr (EXPORT) {
                    // the last row: H(m, c+1) from the lane that holds row m; the last column:
                    // H(i, n) of every row of this lane at column n
                    // Gap halves (affine_dp_algorithm.hpp:139-156): below row m, V(m+1, c+1), and
                    // right of column n, W(i, n+1), the gap states after the last update.  With
                    // the row drift, row r of a lane holds H and W less ge*(r+1), and the vertical
                    // state below it less ge*(r+2).
                    const int m = int(tk.m[0]), n = int(tk.n[0]);
                    if (er >= 0 && er < R && c < n) {
                        int v = 0;
#pragma unroll
                        for (int r = 0; r < R; ++r) if (r == er) v = L::half(st.H[r], 0);
                        dp_row[c + 1] = v + dge * (er + 1);
                        dp_row[a.dp_gap + c + 1] = L::half(vlast, 0) + dge * (er + 2);
                    }
                    if (c == n - 1) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            if (row0 + r + 1 <= m) {
                                dp_col[row0 + r + 1] = L::half(st.H[r], 0) + dge * (r + 1);
                                dp_col[a.dp_gap + row0 + r + 1] = L::half(st.F[r], 0) + dge * (r + 1);
                            }
                        }
                    }
                }
            }
            if constexpr (!LOCAL) {
                // capture events are rare: test the whole wave first
                if (__builtin_expect(__builtin_amdgcn_ballot_w64(k == kev) != 0, 0)) {
                    cap.at_column(c, st, tk, a.scores, dge, -cg * (k % G), -b0, -b1);
                    kev = cap.ev == INT32_MAX ? INT32_MAX : cap.ev + t;
                }
            }
            if constexpr (!ONE) {
                qH = dpp<kRowShl1>(Hout, qH);
                qE = dpp<kRowShl1>(Eout, qE);
            }
            finish_scores<W, R, SCORE>(st, Bnext, tbl_lo, tbl_hi, rd, sc);
            Bin = Bnext;
        };
        auto chunk_end = [&](int k0) {
            conv = nextB;
            nextB = colp[k0 + 2 * G + 1 + t];
            if constexpr (!ONE) {
                convH = nextH;
                convE = nextE;
                top_chunk(k0 + 2 * G + t, nextH, nextE);
            }
            if (has_bnd_out) {
                // lane G-1 computed column k0 + j - (G-1) at the chunk's step j; stored without
                // the column drift it was sent with (cg*(j+1))
                const int col = k0 + t - (G - 1);
                if (col >= 0 && col < N) {
                    uint32_t h = qH, e = qE;
                    if constexpr (CDRIFT) {
                        h = L::bits(L::add(L::from(h), L::splat(-cg * (t + 1))));
                        e = L::bits(L::add(L::from(e), L::splat(-cg * (t + 1))));
                    }
                    bndp[2 * col] = h;
                    bndp[2 * col + 1] = e;
                }
            }
            if constexpr (CDRIFT) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    st.H[r] = L::add(st.H[r], vrebase);
                    st.F[r] = L::add(st.F[r], vrebase);
                }
                Hdiag = L::add(Hdiag, vrebase);
                Hout = L::bits(L::add(L::from(Hout), vrebase));
                Eout = L::bits(L::add(L::from(Eout), vrebase));
                if constexpr (LOCAL) runmax = L::add(runmax, vrebase);
                if constexpr (ONE) {
                    topH = L::bits(L::add(L::from(topH), vrebase));
                    topE = L::bits(L::add(L::from(topE), vrebase));
                }
            }
        };

        // whole chunks of G steps; the first is masked (lanes start one step apart).  Lane G-1
        // computes the last column at step N+G-2: with the byte-permute DNA scores, a stripe
        // with no bottom row to store (the task's last) ends there instead of sweeping pad
        // columns to the end of its last chunk (the other score kinds keep whole chunks: the
        // shorter loop costs their sweeps three VGPRs, which the pair-row table's fused kernel
        // cannot spare at three waves per SIMD)
        const int nsteps = ((N + 2 * G - 2) / G) * G;
        const int last_step = N + G - 2;
        constexpr bool kCutLast = SCORE == kScoreDna;
        constexpr bool kMaskAll = ROWMAX || TRACK || EXPORT;
        auto sweep = [&]<bool TOPSTEP>() {
#pragma unroll 1
            for (int u = 0; u < G; ++u) step.template operator()<true, TOPSTEP>(u);
            chunk_end(0);
            for (int k0 = G; k0 < nsteps; k0 += G) {
                // the last chunk of the task's last stripe ends after the pair of steps that holds
                // the last useful one (k0 <= last_step here)
                const bool cut = kCutLast && !has_bnd_out && k0 + G - 1 > last_step;
                const int pairs = cut ? (last_step - k0 + 2) / 2 : G / 2;
                // two steps per iteration: the register allocator then alternates the exchanged
                // values between two sets of registers instead of copying them back every step
#pragma unroll 1
                for (int p = 0; p < pairs; ++p) {
                    step.template operator()<kMaskAll, TOPSTEP>(k0 + 2 * p);
                    step.template operator()<kMaskAll, TOPSTEP>(k0 + 2 * p + 1);
                }
                if (cut) {
                    // the local running maximum goes back to the frame of step -1 by the steps taken
                    if constexpr (LOCAL && CDRIFT) runmax = L::add(runmax, L::splat(ge * 2 * pairs));
                    break;
                }
                chunk_end(k0);
            }
        };
        if (ONE && top_step != 0) sweep.template operator()<ONE>();
        else sweep.template operator()<false>();
        // free last row only: the int16 saturation tracking (TRACK) keeps per-row maxima too, but
        // they are not candidates for the score
        if constexpr (ROWMAX) cap.row_maximum(st, tk, a.scores, dge);
        fold_rows<W, R, ACC, DRIFT>(st, runmax, satmax, ge);
        if (has_bnd_out) {
            // the boundary row written by lane G-1 is read by lane 0 of the next stripe
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    };
    if (!EXPORT && tk.S == 1) {
        stripe.template operator()<!EXPORT>(0);
    } else {
        for (uint32_t s = 0; s < tk.S; ++s) stripe.template operator()<false>(s);
    }

    if constexpr (LOCAL || TRACK) {
        const V mv = row_reduce_max<W>(runmax);
        const V sv = row_reduce_max<W>(LOCAL ? runmax : satmax);
        if (t == 0) {
#pragma unroll
            for (int q = 0; q < L::NH; ++q) {
                if (tk.pair[q] == kNone) continue;
                // the local maximum is in the frame of step -1 with the column drift
                const int v = L::half(mv, q) + (LOCAL && CDRIFT ? cg : 0) - lbias;
                if constexpr (LOCAL) a.scores[tk.pair[q]] = v;
                // int16 overflow: a value clamped at kMax, or (max3 variants) above kInt16Top,
                // whose NaN pattern the running maximum keeps at 0x7E00 or more; in the frame
                // of step u, so with the local column drift at least G|ge| less in the frame
                // of step -1
                if (W == 16 && L::half(sv, q) >= (MAX3 ? kInt16Top + 1 : L::kMax) - (LOCAL && CDRIFT ? G * cg : 0))
                    atomicOr(a.flags + tk.pair[q], kFlagSaturated);
                if (a.maxh) atomicMax(a.maxh + tk.pair[q], v);
            }
        }
    }
}

// Copies the launch's substitution table into LDS for the score kinds that read one there.
template <int SCORE>
__device__ __forceinline__ void load_launch_table(char *lds_raw, const KernelArgs &a)
{
    if constexpr (SCORE == kScoreMatrix) {
        load_lds_table(lds_raw, a.table, a.table_entries);
        __syncthreads();
    } else if constexpr (SCORE == kScorePairs || SCORE == kScoreProfile) {
        load_lds_pairs(lds_raw, reinterpret_cast<const uint32_t *>(a.table), a.table_entries);
        __syncthreads();
    }
}

// The striped kernel: one task per lane group (one block per 256 / G tasks), or with a.queue only
// as many blocks as the GPU keeps resident, every wave taking the next 64 / G tasks of the launch
// from an atomic counter until none are left (the tasks are in decreasing order of work); the
// last block to finish resets the counter for the next pass.
template <int W, int G, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX, bool EXPORT = false>
__global__ __launch_bounds__(256) void stripe_kernel(KernelArgs a)
{
    extern __shared__ char lds_raw[];
    load_launch_table<SCORE>(lds_raw, a);

    if (!a.queue) {
        const uint32_t gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
        if (gid < a.ntasks) stripe_task<W, G, R, SCORE, LOCAL, TRACK, ROWMAX, EXPORT>(a, gid, lds_raw);
        return;
    }
    const uint32_t group = (threadIdx.x % 64) / G;   // of the wave's 64 / G groups
    for (;;) {
        // the whole wave is here (converged): lane 0 takes the next tasks for it
        uint32_t base = 0;
        if (threadIdx.x % 64 == 0) base = atomicAdd(a.queue, 64u / G);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= a.ntasks) break;
        if (base + group < a.ntasks) stripe_task<W, G, R, SCORE, LOCAL, TRACK, ROWMAX, EXPORT>(a, base + group, lds_raw);
    }
    __syncthreads();   // every wave of the block has seen the queue empty
    if (threadIdx.x == 0 && atomicAdd(a.queue + 1, 1u) == gridDim.x - 1) {
        // the last block: no wave of this launch takes a task any more
        __hip_atomic_store(a.queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.queue + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One persistent launch for the striped launches of a plan that differ only in their rows per
// lane R, within [RLO, RHI] (pa_plan.hpp, FusedLaunch, fused_rows): as many blocks as the GPU
// keeps resident, whose waves take wave items from one atomic counter in the planner's order
// (decreasing modelled time across the member launches).  An item names up to 64 / 16
// consecutive tasks of one member and its R, which selects the sweep; the wave's lane groups take
// one task each.  The last block to finish resets the counter for the next pass.
template <int W, int SCORE, bool LOCAL, int RLO, int RHI>
__global__ __launch_bounds__(256) void fused_stripe_kernel(KernelArgs a)
{
    constexpr int G = 16;
    extern __shared__ char lds_raw[];
    load_launch_table<SCORE>(lds_raw, a);
    const uint32_t group = (threadIdx.x % 64) / G;   // of the wave's 64 / G groups
    // the sweep of rows per lane RR, compiled only inside the kernel's range
    auto sweep = [&]<int RR>(uint32_t gid) {
        if constexpr (RR >= RLO && RR <= RHI) stripe_task<W, G, RR, SCORE, LOCAL, false, false, false>(a, gid, lds_raw);
    };
    for (;;) {
        // the whole wave is here (converged): lane 0 takes the next item for it
        uint32_t idx = 0;
        if (threadIdx.x % 64 == 0) idx = atomicAdd(a.queue, 1u);
        idx = __builtin_amdgcn_readfirstlane(idx);
        if (idx >= a.nitems) break;
        const uint32_t item = __builtin_amdgcn_readfirstlane(a.items[idx]);
        if (group <= ((item >> 27) & 3u)) {
            const uint32_t gid = (item & (kFusedMaxTasks - 1u)) + group;
            switch (item >> 29) {   // (R - 4) / 2, uniform over the wave; the planner keeps R in range
            case 0: sweep.template operator()<4>(gid); break;
            case 1: sweep.template operator()<6>(gid); break;
            case 2: sweep.template operator()<8>(gid); break;
            case 3: sweep.template operator()<10>(gid); break;
            case 4: sweep.template operator()<12>(gid); break;
            case 5: sweep.template operator()<14>(gid); break;
            default: sweep.template operator()<16>(gid); break;
            }
        }
    }
    __syncthreads();   // every wave of the block has seen the queue empty
    if (threadIdx.x == 0 && atomicAdd(a.queue + 1, 1u) == gridDim.x - 1) {
        // the last block: no wave of this launch takes an item any more
        __hip_atomic_store(a.queue, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.queue + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------------------------
// Wavefront kernel for long pairs.
// ---------------------------------------------------------------------------------------------

constexpr int kWfWaves = 4;     // waves per workgroup
constexpr int kWfRing = 256;    // columns per LDS ring between consecutive waves
constexpr int kWfOffSlots = 8;  // chunks of ring offsets kept per wave (the ring spans 4 chunks)
// Bounded waits: after this many polls a wait gives up and flags the batch, so a protocol
// error ends as a reported failure instead of a hang.
constexpr uint32_t kWfSpinLimit = 1u << 24;

struct WfShared {
    uint32_t prod[kWfWaves];   // positions (round * N + column) published by wave w
    uint32_t cons[kWfWaves];   // positions consumed by wave w from its predecessor's ring
    uint32_t failed;           // a wait gave up: later waits return at once
    uint32_t ring[kWfWaves - 1][kWfRing][2];   // (H, V) from wave w to wave w+1
    // int16 re-centring: the offsets of each half of wave w's ring values, per chunk of G
    // columns (chunk c = columns [cG - G + 1, cG], slot c % kWfOffSlots)
    int ring_off[kWfWaves - 1][kWfOffSlots][2];
};

__device__ __forceinline__ uint32_t lds_load_acquire(uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until *p >= need.  After kWfSpinLimit polls it gives up, flags the batch in *err and
// marks the workgroup failed, so that every later wait of the workgroup returns at once and a
// protocol error costs one time-out, not one per wait.
__device__ __forceinline__ void wait_at_least(uint32_t *p, uint32_t need, uint32_t *err, uint32_t *failed)
{
    uint32_t spins = 0;
    while (lds_load_acquire(p) < need) {
        if (lds_load_acquire(failed)) return;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kWfSpinLimit) {
            atomicOr(err, 1u);
            __hip_atomic_store(failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            return;
        }
    }
}

template <int W, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX>
__global__ __launch_bounds__(64 * kWfWaves) void wavefront_kernel(KernelArgs a)
{
    using L = Lane<W>;
    using V = typename L::V;
    constexpr int G = 64;
    constexpr int ROWS = G * R;
    extern __shared__ char lds_raw[];
    __shared__ WfShared sh;

    if (threadIdx.x < kWfWaves) { sh.prod[threadIdx.x] = 0; sh.cons[threadIdx.x] = 0; }
    if (threadIdx.x == 0) sh.failed = 0;
    if constexpr (SCORE == kScoreMatrix) load_lds_table(lds_raw, a.table, a.table_entries);
    if constexpr (SCORE == kScorePairs)
        load_lds_pairs(lds_raw, reinterpret_cast<const uint32_t *>(a.table), a.table_entries);
    __syncthreads();

    const Task tk = a.tasks[blockIdx.x];
    const int wave = int(threadIdx.x / G);
    const int t = int(threadIdx.x % G);
    const uint32_t lane_bank = (threadIdx.x & 31u) * 4u;
    const int go = a.go, ge = a.ge;
    const int fc = LOCAL ? 1 : a.fc, fr = LOCAL ? 1 : a.fr, lc = LOCAL ? 0 : a.lc;
    if (a.self_init) {   // uniform: every wave reaches the barrier
        if (threadIdx.x == 0) init_task_pairs<W, LOCAL>(a, tk, fc, fr, lc, LOCAL ? 0 : a.lr);
        __syncthreads();
    }
    constexpr bool DRIFT = !LOCAL;           // row drift (pa_cells.hpp) for the global method
    constexpr bool ACC = ROWMAX || TRACK;    // per-row maxima
    // column drift: on top of the row drift for the global method, alone for the local method
    constexpr bool CDRIFT = !ACC;
    // int16 global pairs are re-centred: each wave keeps its values relative to a per-half
    // offset (int32, wave-uniform), moved at every chunk end to the value of lane 0's first row,
    // so that the band of cells in flight (within about 1,200 steps of each other) stays far
    // inside the int16 range whatever the absolute scores; the planner uses int16 here only
    // when that spread bound holds.  Values crossing to another wave carry their writer's
    // offsets (the LDS ring per chunk, the HBM row as exact int32).
    constexpr bool RC = W == 16 && CDRIFT && !LOCAL;
    // int16 local lanes take three-way maxima in one instruction, their values biased by the
    // launch's local bias into its range (pa_plan.hpp, kInt16Top); the re-centred lanes keep
    // two-operand maxima
    constexpr bool MAX3 = W == 16 && CDRIFT && LOCAL;
    const int lbias = MAX3 ? a.lbias : 0;
    const int dge = DRIFT ? ge : 0;
    const int cg = CDRIFT ? -ge : 0;         // column drift per step
    const V vgo = L::splat(go), vgogo = L::splat(go + ge), vge = L::splat(ge);
    // shifts the last row's drift back when H, V leave the lane (and, with the column drift,
    // forward by one step: the next lane uses them one step later)
    const V vout = L::splat((DRIFT ? ge * R : 0) + (CDRIFT ? -ge : 0));
    const V vrebase = L::splat(ge * G);      // end of a chunk: the column drift drops by G|ge|
    uint32_t tbl_lo = a.tbl_lo;
    const uint32_t tbl_hi = a.tbl_hi;
    // the permute's second table word must be a VGPR (one scalar operand per instruction): keep
    // it in one for the whole kernel instead of moving it from its scalar register every step
    if constexpr (SCORE == kScoreDna) asm volatile("" : "+v"(tbl_lo));
    const int N = int(tk.N);
    const int S = int(tk.S);
    const uint32_t *colp = a.colcodes + tk.col_off;
    uint32_t *gbnd = a.bnd + 2ull * tk.bnd_off;   // last wave -> first wave, one full row
    // (re-centred: four words per column, the exact H and V of each half)

    // exact; for the local method (column drift) 0 in the frame of step -1, moved into each
    // step's frame as the step begins and back by G steps at each chunk end
    V runmax = LOCAL ? L::splat(-cg + lbias) : L::lowest();
    V satmax = L::lowest();                         // drifted (int16 saturation test)
    const V vcg = L::splat(cg);

    for (int s = wave; s < S; s += kWfWaves) {
        const int round = s / kWfWaves;
        const uint32_t pos0 = uint32_t(round) * uint32_t(N);
        const int row0 = s * ROWS + t * R;
        Rows<W, R, ACC> st;
        // re-centring offsets of the two halves: the exact H(first row of the stripe, 0) to start
        int off0 = 0, off1 = 0;
        if constexpr (RC) off0 = off1 = h0_value(s * ROWS, fc, go, ge);
        // lane t starts at step t: its initial column (j = 0) counts as computed at step t-1
        init_rows<W, R, SCORE, ACC, DRIFT, CDRIFT>(st, a.rowcodes + tk.row_off, row0, fc, go, ge, lane_bank,
                                                   cg * (t - 1) - off0 + lbias);
        V Hdiag = L::splat(h0_value(row0, fc, go, ge) + cg * (t - 1) - off0 + lbias);
        Capture<W, R> cap;
        if constexpr (!LOCAL) cap.init(tk, row0, lc);

        // top boundary source: 0 = initial row, 1 = LDS ring of wave-1, 2 = HBM row of the last wave
        const int src = s == 0 ? 0 : (wave > 0 ? 1 : 2);
        // bottom boundary sink: 0 = none (last stripe), 1 = LDS ring of this wave, 2 = HBM row
        const int dst = s + 1 >= S ? 0 : (wave < kWfWaves - 1 ? 1 : 2);
        uint32_t *src_prog = src == 1 ? &sh.prod[wave - 1] : &sh.prod[kWfWaves - 1];
        const uint32_t src_pos0 = src == 1 ? pos0 : pos0 - uint32_t(N);   // HBM row: previous round

        auto need_top = [&](int upto) {   // columns [0, upto) of the top boundary are published
            if (src == 0) return;
            wait_at_least(src_prog, src_pos0 + uint32_t(upto < N ? upto : N), a.flags + a.error_slot, &sh.failed);
            if (src == 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        };
        // lane t of a chunk is consumed at the chunk's step t: column drift cg*t
        auto read_top = [&](int col, uint32_t &h, uint32_t &e) {
            if (src == 1) {
                const uint32_t slot = (pos0 + uint32_t(col)) % kWfRing;
                h = sh.ring[wave - 1][slot][0];
                e = sh.ring[wave - 1][slot][1];
                if constexpr (RC) {   // from the writer's offsets of that chunk to ours
                    const int c = (col + G - 1) / G;
                    const int *wo = sh.ring_off[wave - 1][c % kWfOffSlots];
                    const V d = L::pair(wo[0] - off0, wo[1] - off1);
                    h = L::bits(L::add(L::from(h), d));
                    e = L::bits(L::add(L::from(e), d));
                }
            } else if constexpr (RC) {   // exact values of each half
                const int *x = reinterpret_cast<const int *>(gbnd) + 4 * col;
                h = L::bits(L::pair(x[0] - off0, x[1] - off1));
                e = L::bits(L::pair(x[2] - off0, x[3] - off1));
            } else {
                h = gbnd[2 * col];
                e = gbnd[2 * col + 1];
            }
            if constexpr (CDRIFT) {
                h = L::bits(L::add(L::from(h), L::splat(cg * t)));
                e = L::bits(L::add(L::from(e), L::splat(cg * t)));
            }
        };
        auto publish_consumed = [&](int upto) {
            if (src == 1 && t == 0)
                __hip_atomic_store(&sh.cons[wave], pos0 + uint32_t(upto < N ? upto : N), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        };

        // Conveyors as in the striped kernel, one column per lane: chunk j of the column-code
        // conveyor holds columns jG+1 .. jG+G, the top-boundary conveyor columns jG .. jG+G-1.
        uint32_t Bin = colp[t];
        uint32_t conv = colp[1 + t];
        uint32_t nextB = colp[G + 1 + t];
        uint32_t convH = 0, convE = 0, nextH = 0, nextE = 0;
        if (src != 0) {
            need_top(2 * G);
            if (t < N) read_top(t, convH, convE);
            if (G + t < N) read_top(G + t, nextH, nextE);
            publish_consumed(2 * G);
        }
        uint32_t Hout = 0, Eout = 0;
        uint32_t qH = 0, qE = 0;   // bottom-row outputs of lane G-1 over one chunk
        int hb_top = h0_value(1, fr, go, ge);

        ScoreReads<W, R, SCORE> rd;
        issue_scores<W, R, SCORE>(st, Bin, lds_raw, rd);
        V sc[R];
        finish_scores<W, R, SCORE>(st, Bin, tbl_lo, tbl_hi, rd, sc);

        const int nsteps = ((N + 2 * G - 2) / G) * G;
        for (int k0 = 0; k0 < nsteps; k0 += G) {
#pragma unroll 2
            for (int u = 0; u < G; ++u) {
                const int c = k0 + u - t;
                const uint32_t shifted = dpp_any<kWaveShl1>(conv);
                const uint32_t Bnext = dpp<kWaveShr1>(conv, Bin);
                conv = shifted;
                issue_scores<W, R, SCORE>(st, Bnext, lds_raw, rd);
                uint32_t topH, topE;
                if (src != 0) {
                    topH = convH;
                    topE = convE;
                } else if constexpr (RC) {
                    const int h = hb_top + cg * u, e = hb_top + go + ge - dge + cg * u;
                    topH = L::bits(L::pair(h - off0, h - off1));
                    topE = L::bits(L::pair(e - off0, e - off1));
                    hb_top = fr ? 0 : hb_top + ge;
                } else {
                    topH = L::top_word(hb_top + cg * u + lbias);
                    topE = L::top_word(hb_top + go + ge - dge + cg * u + lbias);   // drifted by |ge| for row 0
                    hb_top = fr ? 0 : hb_top + ge;
                }
                if (src != 0) {
                    convH = dpp_any<kWaveShl1>(convH);
                    convE = dpp_any<kWaveShl1>(convE);
                }
                const uint32_t Hin = dpp<kWaveShr1>(topH, Hout);
                const uint32_t Ein = dpp<kWaveShr1>(topE, Eout);
                if constexpr (LOCAL) runmax = L::add(runmax, vcg);   // into this step's frame
                if (uint32_t(c) < uint32_t(N)) {
                    V Eup = L::from(Ein);
                    update_cells<W, R, SCORE, LOCAL, ACC, DRIFT, CDRIFT, MAX3>(st, Eup, Hdiag, runmax, sc, Bin, tbl_lo,
                                                                               tbl_hi, vgo, vgogo, vge,
                                                                               L::splat(cg * u + lbias));
                    Hdiag = L::from(Hin);
                    Hout = L::bits(DRIFT || CDRIFT ? L::add(st.H[R - 1], vout) : st.H[R - 1]);
                    Eout = L::bits(DRIFT || CDRIFT ? L::add(Eup, vout) : Eup);
                    if constexpr (!LOCAL) cap.at_column(c, st, tk, a.scores, dge, -cg * u, off0, off1);
                }
                if (dst != 0) {
                    qH = dpp<kWaveShl1>(Hout, qH);
                    qE = dpp<kWaveShl1>(Eout, qE);
                }
                finish_scores<W, R, SCORE>(st, Bnext, tbl_lo, tbl_hi, rd, sc);
                Bin = Bnext;
            }
            // end of chunk: refill the conveyors one chunk ahead
            conv = nextB;
            nextB = colp[k0 + 2 * G + 1 + t];
            if (src != 0) {
                convH = nextH;
                convE = nextE;
                need_top(k0 + 3 * G);
                if (k0 + 2 * G + t < N) read_top(k0 + 2 * G + t, nextH, nextE);
                publish_consumed(k0 + 3 * G);
            }
            if (dst != 0) {
                // lane j holds the bottom row of column k0 + j - (G-1), sent with column drift
                // cg*(j+1): stored without it
                const int col = k0 + t - (G - 1);
                const int hi = k0 + 1 < N ? k0 + 1 : N;   // columns [0, hi) done after this chunk
                if constexpr (CDRIFT) {
                    qH = L::bits(L::add(L::from(qH), L::splat(-cg * (t + 1))));
                    qE = L::bits(L::add(L::from(qE), L::splat(-cg * (t + 1))));
                }
                if (dst == 1) {
                    // the slots of [k0+1-G, k0+1) held positions kWfRing earlier: those must
                    // have been consumed by the next wave
                    const int64_t need = int64_t(pos0) + hi - kWfRing;
                    if (need > 0) wait_at_least(&sh.cons[wave + 1], uint32_t(need), a.flags + a.error_slot, &sh.failed);
                    if (col >= 0 && col < N) {
                        const uint32_t slot = (pos0 + uint32_t(col)) % kWfRing;
                        sh.ring[wave][slot][0] = qH;
                        sh.ring[wave][slot][1] = qE;
                    }
                    if constexpr (RC) {   // this chunk's offsets, published with its columns
                        if (t == 0) {
                            sh.ring_off[wave][(k0 / G) % kWfOffSlots][0] = off0;
                            sh.ring_off[wave][(k0 / G) % kWfOffSlots][1] = off1;
                        }
                    }
                } else if (col >= 0 && col < N) {
                    if constexpr (RC) {
                        int *x = reinterpret_cast<int *>(gbnd) + 4 * col;
                        x[0] = L::half(L::from(qH), 0) + off0;
                        x[1] = L::half(L::from(qH), 1) + off1;
                        x[2] = L::half(L::from(qE), 0) + off0;
                        x[3] = L::half(L::from(qE), 1) + off1;
                    } else {
                        gbnd[2 * col] = qH;
                        gbnd[2 * col + 1] = qE;
                    }
                }
                if (hi > k0 + 1 - G) {   // this chunk produced columns: publish them
                    if (dst == 2) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                    if (t == G - 1)
                        __hip_atomic_store(&sh.prod[wave], pos0 + uint32_t(hi), __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            if constexpr (RC) {
                // re-centre on lane 0's first row, then drop the column drift as below
                const uint32_t ref = __builtin_amdgcn_readlane(L::bits(st.H[0]), 0);
                const int d0 = int(int16_t(ref & 0xFFFFu)), d1 = int(int16_t(ref >> 16));
                off0 += d0;
                off1 += d1;
                const V vd = L::pair(-d0, -d1);
                const V vt = L::pair(ge * G - d0, ge * G - d1);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    st.H[r] = L::add(st.H[r], vt);
                    st.F[r] = L::add(st.F[r], vt);
                }
                Hdiag = L::add(Hdiag, vt);
                Hout = L::bits(L::add(L::from(Hout), vt));
                Eout = L::bits(L::add(L::from(Eout), vt));
                // top-boundary values already converted to this wave's offsets
                convH = L::bits(L::add(L::from(convH), vd));
                convE = L::bits(L::add(L::from(convE), vd));
                nextH = L::bits(L::add(L::from(nextH), vd));
                nextE = L::bits(L::add(L::from(nextE), vd));
            } else if constexpr (CDRIFT) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    st.H[r] = L::add(st.H[r], vrebase);
                    st.F[r] = L::add(st.F[r], vrebase);
                }
                Hdiag = L::add(Hdiag, vrebase);
                Hout = L::bits(L::add(L::from(Hout), vrebase));
                Eout = L::bits(L::add(L::from(Eout), vrebase));
                if constexpr (LOCAL) runmax = L::add(runmax, vrebase);
            }
        }
        // free last row only: the int16 saturation tracking (TRACK) keeps per-row maxima too, but
        // they are not candidates for the score
        if constexpr (ROWMAX) cap.row_maximum(st, tk, a.scores, dge);
        fold_rows<W, R, ACC, DRIFT>(st, runmax, satmax, ge);
    }

    if constexpr (LOCAL || TRACK) {
        const V mv = row_reduce_max<W>(runmax);
        const V sv = row_reduce_max<W>(LOCAL ? runmax : satmax);
        if ((t & 15) == 0) {
#pragma unroll
            for (int q = 0; q < L::NH; ++q) {
                if (tk.pair[q] == kNone) continue;
                // the local maximum is in the frame of step -1 (column drift)
                const int v = L::half(mv, q) + (LOCAL ? cg : 0) - lbias;
                if constexpr (LOCAL) atomicMax(a.scores + tk.pair[q], v);
                // int16 overflow (a value clamped at kMax, or above kInt16Top in the max3 lanes),
                // in the frame of step u: at least G|ge| less in the frame of step -1
                if (W == 16 && L::half(sv, q) >= (MAX3 ? kInt16Top + 1 : L::kMax) - (LOCAL && CDRIFT ? G * cg : 0))
                    atomicOr(a.flags + tk.pair[q], kFlagSaturated);
                if (a.maxh) atomicMax(a.maxh + tk.pair[q], v);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Launch dispatch
// ---------------------------------------------------------------------------------------------

// The grid of a striped launch: one block per 256 / 16 tasks, or with the work queue no more
// blocks than the GPU keeps resident (its CUs times the kernel's blocks per CU).
template <typename K>
static dim3 striped_grid(K kernel, const KernelArgs &a, uint32_t blocks, uint32_t lds)
{
    if (!a.queue) return dim3(blocks);
    int per_cu = 0, device = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 1;
    uint32_t resident = uint32_t(per_cu) * uint32_t(cus);
    // a fused launch that shares its pass with other launches takes its share of the GPU
    if (a.items && a.share > 0.0f && a.share < 1.0f)
        resident = std::max<uint32_t>(uint32_t(cus), uint32_t(double(resident) * a.share + 0.5));
    return dim3(std::min<uint32_t>(blocks, resident));
}

template <int W, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX>
static void launch_striped(const VariantKey &, const KernelArgs &a, dim3 blocks, dim3 threads, uint32_t lds,
                           hipStream_t st)
{
    auto kernel = stripe_kernel<W, 16, R, SCORE, LOCAL, TRACK, ROWMAX>;
    hipLaunchKernelGGL(kernel, striped_grid(kernel, a, blocks.x, lds), threads, lds, st, a);
}

template <int W, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX, bool EXPORT = false>
static hipError_t launch_variant(const VariantKey &k, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st)
{
    const uint32_t lds = SCORE != kScoreDna ? lds_bytes : 0;
    const uint32_t threads = 256;
    const uint32_t blocks = uint32_t((uint64_t(a.ntasks) * 16 + threads - 1) / threads);
    if constexpr (EXPORT) {
        if (k.G != 16) return hipErrorInvalidValue;
        auto kernel = stripe_kernel<W, 16, R, SCORE, LOCAL, TRACK, ROWMAX, true>;
        hipLaunchKernelGGL(kernel, striped_grid(kernel, a, blocks, lds), dim3(threads), lds, st, a);
        return hipGetLastError();
    } else if constexpr (SCORE == kScoreProfile) {
        // striped kernel only, at rows per lane whose profile reads are free of bank conflicts
        if constexpr (R == 8 || R == 16) {
            return hipErrorInvalidValue;
        } else {
            if (k.G != 16) return hipErrorInvalidValue;
            launch_striped<W, R, SCORE, LOCAL, TRACK, ROWMAX>(k, a, dim3(blocks), dim3(threads), lds, st);
            return hipGetLastError();
        }
    } else {
        if (k.G == 64) {
            hipLaunchKernelGGL((wavefront_kernel<W, R, SCORE, LOCAL, TRACK, ROWMAX>), dim3(a.ntasks),
                               dim3(64 * kWfWaves), lds, st, a);
            return hipGetLastError();
        }
        launch_striped<W, R, SCORE, LOCAL, TRACK, ROWMAX>(k, a, dim3(blocks), dim3(threads), lds, st);
        return hipGetLastError();
    }
}

template <int W, int R, int SCORE, bool LOCAL>
static hipError_t dispatch_flags(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    if (k.export_dp) {
        if constexpr (W == 32 && R == kExportRows && SCORE != kScoreProfile && SCORE != kScorePairs) {
            if (LOCAL) return launch_variant<W, R, SCORE, LOCAL, false, false, true>(k, a, lds, st);
            if (k.rowmax) return launch_variant<W, R, SCORE, LOCAL, false, true, true>(k, a, lds, st);
            return launch_variant<W, R, SCORE, LOCAL, false, false, true>(k, a, lds, st);
        } else {
            return hipErrorInvalidValue;
        }
    }
    if constexpr (LOCAL) {   // local variants keep no per-row maxima
        return launch_variant<W, R, SCORE, LOCAL, false, false>(k, a, lds, st);
    } else {
        if (k.rowmax) {
            if (k.track) return launch_variant<W, R, SCORE, LOCAL, true, true>(k, a, lds, st);
            return launch_variant<W, R, SCORE, LOCAL, false, true>(k, a, lds, st);
        }
        if (k.track) return launch_variant<W, R, SCORE, LOCAL, true, false>(k, a, lds, st);
        return launch_variant<W, R, SCORE, LOCAL, false, false>(k, a, lds, st);
    }
}

template <int W, int R>
hipError_t dispatch_score(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    if (k.score == kScoreDna) {
        return k.local ? dispatch_flags<W, R, kScoreDna, true>(k, a, lds, st)
                       : dispatch_flags<W, R, kScoreDna, false>(k, a, lds, st);
    }
    if constexpr (W == 16) {
        if (k.score == kScorePairs)
            return k.local ? dispatch_flags<W, R, kScorePairs, true>(k, a, lds, st)
                           : dispatch_flags<W, R, kScorePairs, false>(k, a, lds, st);
        if (k.score == kScoreProfile)
            return k.local ? dispatch_flags<W, R, kScoreProfile, true>(k, a, lds, st)
                           : dispatch_flags<W, R, kScoreProfile, false>(k, a, lds, st);
    } else {
        if (k.score == kScoreProfile) return hipErrorInvalidValue;
    }
    return k.local ? dispatch_flags<W, R, kScoreMatrix, true>(k, a, lds, st)
                   : dispatch_flags<W, R, kScoreMatrix, false>(k, a, lds, st);
}

} // namespace pa
