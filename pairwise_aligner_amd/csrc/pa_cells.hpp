// pa_cells.hpp -- device building blocks shared by the striped and the wavefront kernels:
// lane types (packed saturating int16 / int32), DPP helpers, the R-row cell update and the
// score-capture bookkeeping.
//
// Cell recurrence (rrahn/pairwise_aligner affine/affine_dp_algorithm.hpp:139-156):
//   best = H(i-1,j-1) + s(a_i, b_j)         [local: max(best, 0), dp_matrix_local.hpp:39-47]
//   best = max(best, V(i,j), W(i,j))         H(i,j) = best
//   open = best + go + ge
//   V(i+1,j) = max(V(i,j) + ge, open)        vertical gap state, flows down the rows
//   W(i,j+1) = max(W(i,j) + ge, open)        horizontal gap state, kept per row in registers
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_plan.hpp"

namespace pa {

typedef short v2i16 __attribute__((ext_vector_type(2)));

template <int W> struct Lane;

template <> struct Lane<16> {
    typedef v2i16 V;
    static constexpr int NH = 2;
    static constexpr int kMax = 32767;
    __device__ static int clamp(int x) { return x < -32768 ? -32768 : (x > 32767 ? 32767 : x); }
    __device__ static V splat(int x) { short s = short(clamp(x)); return V{s, s}; }
    __device__ static V add(V a, V b) { return __builtin_elementwise_add_sat(a, b); }
    __device__ static V vmax(V a, V b) { return __builtin_elementwise_max(a, b); }
    __device__ static uint32_t bits(V v) { return __builtin_bit_cast(uint32_t, v); }
    __device__ static V from(uint32_t u) { return __builtin_bit_cast(V, u); }
    __device__ static int half(V v, int q) { return q == 0 ? int(v.x) : int(v.y); }
    __device__ static V lowest() { return V{short(-32768), short(-32768)}; }
    __device__ static V pair(int a, int b) { return V{short(clamp(a)), short(clamp(b))}; }   // halves a, b
    // packed top-boundary word of an int value (<= 0): both halves, clamped at -32768
    __device__ static uint32_t top_word(int x) { return (uint32_t(x < -32768 ? -32768 : x) & 0xFFFFu) * 0x10001u; }
    // Maximum of three values, per half, in one instruction: the packed fp16 three-input maximum
    // (v_pk_maximum3_f16).  It is the integer maximum when one operand is in [0, kInt16Top] and
    // each other one is in [0, kInt16Top] or at most -1024: the patterns 0..0x7C00 order as
    // their integers, and 0x8000..0xFC00 are -0, the negative numbers and -infinity, below them
    // (pa_plan.hpp, kInt16Top).
    __device__ static V max3(V a, V b, V c)
    {
        typedef _Float16 H2 __attribute__((ext_vector_type(2)));
        const H2 r = __builtin_elementwise_maximum(
            __builtin_elementwise_maximum(__builtin_bit_cast(H2, a), __builtin_bit_cast(H2, b)), __builtin_bit_cast(H2, c));
        return __builtin_bit_cast(V, r);
    }
};

template <> struct Lane<32> {
    typedef int V;
    static constexpr int NH = 1;
    static constexpr int kMax = INT32_MAX;
    __device__ static V splat(int x) { return x; }
    __device__ static V add(V a, V b) { return a + b; }
    __device__ static V vmax(V a, V b) { return a > b ? a : b; }
    __device__ static uint32_t bits(V v) { return uint32_t(v); }
    __device__ static V from(uint32_t u) { return int(u); }
    __device__ static int half(V v, int) { return v; }
    __device__ static V lowest() { return INT32_MIN; }
    __device__ static V pair(int a, int) { return a; }
    __device__ static uint32_t top_word(int x) { return uint32_t(x); }
    __device__ static V max3(V a, V b, V c) { return vmax(vmax(a, b), c); }
};

// DPP controls (gfx9).  Lanes without a source keep `old` (bound_ctrl off).
constexpr int kRowShr1 = 0x111;   // lane t-1 -> t inside a row of 16 lanes
constexpr int kRowShl1 = 0x101;   // lane t+1 -> t inside a row of 16 lanes
constexpr int kRowRor1 = 0x121;
constexpr int kRowRor2 = 0x122;
constexpr int kRowRor4 = 0x124;
constexpr int kRowRor8 = 0x128;
constexpr int kWaveShr1 = 0x138;  // lane t-1 -> t across the wave
constexpr int kWaveShl1 = 0x130;  // lane t+1 -> t across the wave

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(src), CTRL, 0xF, 0xF, false));
}

// DPP move whose lanes without a source take an unspecified value: for a conveyor shifted towards
// lane 0, whose vacated lane is refilled (at the chunk end) before anything reads it.  Leaves the
// register allocator free to place the result (no copy of a tied old value).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_any(uint32_t src)
{
    return uint32_t(__builtin_amdgcn_mov_dpp(int(src), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ int h0_value(int i, int free_flag, int go, int ge)
{
    return (i == 0 || free_flag) ? 0 : go + ge * i;
}

// Row drift (global method).  Inside a lane, register row r (r = 0..R-1) keeps H and both gap
// states shifted up by |ge|*(r+1):  x' = x - ge*(r+1).  The vertical gap state then needs one
// max per row,
//   V'(r+1) = max(V'(r), max(D', W')(r) + go)        (V(r+1) = max(V(r)+ge, H(r)+go+ge)),
// the diagonal term gains |ge| (D' = H'(r-1) + s - ge, folded into the substitution scores by
// the planner), and values leaving the lane are shifted back: H exact, V + |ge| (the drift of
// the next lane's first row).  Per-row maxima are kept drifted and shifted back when read.  The
// local method has no row drift: its zero clamp would need a different bound per row.
//
// Column drift (global method without per-row maxima, on top of the row drift; local method
// without per-row maxima or the DP export, alone, with the zero clamp raised to |ge|*u and the
// running maximum moved into each step's frame).  A value computed at step k of a chunk of G
// steps (u = k mod G) also carries |ge|*u.  Consecutive
// columns of a row are consecutive steps, so the horizontal gap state loses its extension term
// too, W'(j+1) = max(W'(j), H'(j) + go), and the diagonal term gains another |ge| (scores shifted
// by -2ge).  The drift is the same for every lane at a step, so values passed between lanes are
// shifted by a constant; at the end of a chunk every live value drops by G*|ge|, which keeps the
// drift below |ge|*(R+1+G).
template <bool DRIFT>
__device__ __forceinline__ int drift_of(int ge, int r)
{
    return DRIFT ? -ge * (r + 1) : 0;
}

// Register state of one lane: R consecutive DP rows of the current stripe.  ACC: per-row
// running maxima (local / tracked running maximum, free last row).
template <int W, int R, bool ACC>
struct Rows {
    using V = typename Lane<W>::V;
    uint32_t A[R];        // residue codes (matrix: LDS byte offsets of the table rows)
    uint32_t Ahi[R];      // int16 matrix path: codes of the high half
    V H[R], F[R];
    V rowacc[ACC ? R : 1];
};

// Loads the residue codes of rows row0+1 .. row0+R and sets H(i,0) and W(i,1) (drifted if DRIFT;
// with the column drift CDRIFT, H carries `hcol` more and W enters one column later; values of the
// max3 variants carry their bias `vbias`).
template <int W, int R, int SCORE, bool ACC, bool DRIFT, bool CDRIFT>
__device__ __forceinline__ void init_rows(Rows<W, R, ACC> &st, const uint32_t *codes, int row0, int fc, int go,
                                          int ge, uint32_t lane_bank, int hcol,
                                          typename Lane<W>::V vbias = typename Lane<W>::V{})
{
    using L = Lane<W>;
    const typename L::V vgogo = L::splat(CDRIFT ? go : go + ge);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if constexpr (SCORE == kScorePairs) {
            // one word per two rows: the pair-row table offsets of both halves
            if (r % 2 == 0) st.A[r / 2] = codes[row0 / 2 + r / 2];
            st.Ahi[r] = 0;
        } else if constexpr (SCORE == kScoreProfile) {
            // every task's rows are the query: the profile column of row i is at byte 2i
            (void)codes;
            if (r == 0) st.A[0] = uint32_t(row0) * 2u;
            st.Ahi[r] = 0;
        } else {
            const uint32_t code = codes[row0 + r];
            if constexpr (SCORE == kScoreMatrix && W == 16) {
                st.A[r] = (code & 0xFFFFu) + lane_bank;
                st.Ahi[r] = (code >> 16) + lane_bank;
            } else if constexpr (SCORE == kScoreMatrix) {
                st.A[r] = code + lane_bank;
                st.Ahi[r] = 0;
            } else {
                st.A[r] = code;
                st.Ahi[r] = 0;
            }
        }
        st.H[r] = L::add(L::splat(h0_value(row0 + r + 1, fc, go, ge) + drift_of<DRIFT>(ge, r) + (CDRIFT ? hcol : 0)),
                         vbias);
        st.F[r] = L::add(st.H[r], vgogo);
        if constexpr (ACC) st.rowacc[r] = L::lowest();
    }
}

// Folds the per-row maxima of a stripe into the exact running maximum `runmax` and the drifted
// maximum `satmax` (an int16 saturation test: every drifted value is >= its exact value).
template <int W, int R, bool ACC, bool DRIFT>
__device__ __forceinline__ void fold_rows(const Rows<W, R, ACC> &st, typename Lane<W>::V &runmax,
                                          typename Lane<W>::V &satmax, int ge)
{
    using L = Lane<W>;
    if constexpr (ACC) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            runmax = L::vmax(runmax, L::add(st.rowacc[r], L::splat(-drift_of<DRIFT>(ge, r))));
            satmax = L::vmax(satmax, st.rowacc[r]);
        }
    }
}

// Substitution scores of one lane's R rows against column residue `Bin`, packed like H.
//   DNA codes: v_perm_b32 over an 8-byte table selected by row code ^ column code.
//   matrix: reads of the lane's bank-swizzled LDS table copy at row offset + column offset;
//           split in two so the reads of step k+1 can be issued before the cells of step k.
//   pair-row table / query profile (int16): one dword per two rows and half, interleaved by
//           byte permutes.
template <int W, int R, int SCORE>
struct ScoreReads {
    static constexpr bool kDwordPairs = SCORE == kScorePairs || SCORE == kScoreProfile;
    uint32_t lo[SCORE == kScoreMatrix ? R : (kDwordPairs ? R / 2 : 1)];
    uint32_t hi[SCORE == kScoreMatrix && W == 16 ? R : (kDwordPairs ? R / 2 : 1)];
};

template <int W, int R, int SCORE, bool ACC>
__device__ __forceinline__ void issue_scores(const Rows<W, R, ACC> &st, uint32_t Bin, const char *lds,
                                             ScoreReads<W, R, SCORE> &rd)
{
    if constexpr (SCORE == kScorePairs) {
        // rows 2p and 2p+1 of each half in one dword read
#pragma unroll
        for (int p = 0; p < R / 2; ++p) {
            rd.lo[p] = *reinterpret_cast<const uint32_t *>(lds + (st.A[p] & 0xFFFFu) + (Bin & 0xFFFFu));
            rd.hi[p] = *reinterpret_cast<const uint32_t *>(lds + (st.A[p] >> 16) + (Bin >> 16));
        }
    } else if constexpr (SCORE == kScoreProfile) {
        // the lane's consecutive rows are consecutive int16 entries of the profile column of
        // each half's residue: one base address per half, then immediate offsets
        static_assert(W == 16, "the query profile is an int16 path");
        const char *plo = lds + (Bin & 0xFFFFu) + st.A[0];
        const char *phi = lds + (Bin >> 16) + st.A[0];
#pragma unroll
        for (int p = 0; p < R / 2; ++p) {
            rd.lo[p] = *reinterpret_cast<const uint32_t *>(plo + 4 * p);
            rd.hi[p] = *reinterpret_cast<const uint32_t *>(phi + 4 * p);
        }
    } else if constexpr (SCORE == kScoreMatrix) {
        uint32_t Blo = Bin, Bhi = 0;
        if constexpr (W == 16) { Blo = Bin & 0xFFFFu; Bhi = Bin >> 16; }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (W == 16) {
                rd.lo[r] = *reinterpret_cast<const unsigned short *>(lds + st.A[r] + Blo);
                rd.hi[r] = *reinterpret_cast<const unsigned short *>(lds + st.Ahi[r] + Bhi);
            } else {
                rd.lo[r] = uint32_t(int(*reinterpret_cast<const short *>(lds + st.A[r] + Blo)));
            }
        }
    }
}

template <int W, int R, int SCORE, bool ACC>
__device__ __forceinline__ void finish_scores(const Rows<W, R, ACC> &st, uint32_t Bin, uint32_t tbl_lo,
                                              uint32_t tbl_hi, const ScoreReads<W, R, SCORE> &rd,
                                              typename Lane<W>::V (&sc)[R])
{
    using L = Lane<W>;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if constexpr (SCORE == kScoreDna) {
            // formed where it is used (update_cells): two VALU instructions, nothing to hide
            (void)tbl_lo; (void)tbl_hi; (void)Bin;
        } else if constexpr (SCORE == kScorePairs || SCORE == kScoreProfile) {
            // (low halves) = row 2p of both pairs, (high halves) = row 2p+1
            sc[r] = L::from(__builtin_amdgcn_perm(rd.hi[r / 2], rd.lo[r / 2], (r % 2) ? 0x07060302u : 0x05040100u));
        } else if constexpr (W == 16) {
            sc[r] = L::from((rd.hi[r] << 16) | rd.lo[r]);
        } else {
            sc[r] = L::from(rd.lo[r]);
        }
    }
}

// The R cells of DP column j for one lane's rows, given their substitution scores.  DRIFT: the
// rows are drifted (see drift_of) and the scores already shifted by |ge|; `Eup` enters as the
// vertical gap state of row 0 and leaves as that of row R.  `diag` is the exact H of the row above
// at column j-1.  RUN: keep the exact running maximum in `runmax` (local method); per-row
// maxima go to rowacc (ACC).  MAX3 (int16 lanes with the column drift, values biased into the
// three-way maximum's range): the local clamp and max(D, W) are one three-way maximum, the
// running maximum takes two rows per instruction, and the global method's max(D, W, V) is one.
// CAPV (the DP export, no column drift): *vcap receives the vertical gap state below row vrow
// of this lane (Eup after that row; unchanged when vrow is not one of the lane's rows).
template <int W, int R, int SCORE, bool LOCAL, bool ACC, bool DRIFT, bool CDRIFT, bool MAX3 = false,
          bool CAPV = false>
__device__ __forceinline__ void update_cells(Rows<W, R, ACC> &st, typename Lane<W>::V &Eup,
                                             typename Lane<W>::V diag, typename Lane<W>::V &runmax,
                                             const typename Lane<W>::V (&sc)[R], uint32_t Bin, uint32_t tbl_lo,
                                             uint32_t tbl_hi, typename Lane<W>::V vgo,
                                             typename Lane<W>::V vgogo, typename Lane<W>::V vge,
                                             typename Lane<W>::V vfloor = typename Lane<W>::V{},
                                             typename Lane<W>::V *vcap = nullptr, int vrow = -1)
{
    using L = Lane<W>;
    using V = typename L::V;
    static_assert(!MAX3 || (W == 16 && CDRIFT), "three-way maxima: int16 lanes with the column drift");
    static_assert(!CAPV || !CDRIFT, "the gap-state capture serves the DP export (no column drift)");
    // Diagonal terms of all R rows first: they read only the previous column's H, so every
    // H[r] can then be overwritten in place (no register rotation from step to step).
    V D[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        V s;
        if constexpr (SCORE == kScoreDna) s = L::from(__builtin_amdgcn_perm(tbl_hi, tbl_lo, st.A[r] ^ Bin));
        else s = sc[r];
        D[r] = L::add(r == 0 ? diag : st.H[r - 1], s);
        // the local clamp at 0; with the column drift, 0 of this step's frame (vfloor)
        if constexpr (LOCAL && !MAX3) D[r] = L::vmax(D[r], CDRIFT ? vfloor : L::splat(0));
    }
    V DF[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if constexpr (MAX3 && LOCAL) DF[r] = L::max3(D[r], vfloor, st.F[r]);   // the clamp included
        else if constexpr (!MAX3) DF[r] = L::vmax(D[r], st.F[r]);
    }
    if constexpr (CDRIFT && !DRIFT) {
        // Column drift only (local method): every value of a step carries the step's drift, so
        // the horizontal gap state loses its extension term, W'(j+1) = max(W'(j), H'(j) + go),
        // and the vertical one keeps it: V(r+1) = max(V(r) + ge, H(r) + go + ge)
        // = max(V(r), H(r) + go) + ge, sharing the opening term H + go with W.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const V best = L::vmax(DF[r], Eup);
            const V open = L::add(best, vgo);
            Eup = L::add(L::vmax(Eup, open), vge);
            st.H[r] = best;
            if constexpr (LOCAL && MAX3) {   // rows r-1 and r in one three-way maximum
                if (r % 2 == 1) runmax = L::max3(runmax, st.H[r - 1], best);
                else if (r == R - 1) runmax = L::vmax(runmax, best);
            } else if constexpr (LOCAL) {
                runmax = L::vmax(runmax, best);
            }
            st.F[r] = L::vmax(st.F[r], open);
        }
    } else if constexpr (CDRIFT) {
        // Row and column drift: both gap states lose their extension term,
        //   V'(r+1) = max(V'(r), H'(r) + go),   W'(j+1) = max(W'(j), H'(j) + go),
        // and share the opening term.
#pragma unroll
        for (int r = 0; r < R; ++r) {
            V best;
            if constexpr (MAX3) best = L::max3(D[r], st.F[r], Eup);
            else best = L::vmax(DF[r], Eup);
            const V open = L::add(best, vgo);
            Eup = L::vmax(Eup, open);
            st.H[r] = best;
            if constexpr (LOCAL) runmax = L::vmax(runmax, best);
            if constexpr (ACC) st.rowacc[r] = L::vmax(st.rowacc[r], best);
            st.F[r] = L::vmax(st.F[r], open);
        }
    } else if constexpr (DRIFT) {
        // With go <= 0 (so go+ge <= ge) the vertical gap state needs only max(D, W) of its row:
        //   V(i+1,j) = max(V + ge, H + go + ge) = max(V + ge, max(D, W) + go + ge)
        // because the dropped term V + go + ge never exceeds V + ge; drifted, V' = max(V', DF' + go):
        // one max on the row-to-row chain.
        V DFo[R];
#pragma unroll
        for (int r = 0; r < R; ++r) DFo[r] = L::add(DF[r], vgo);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const V best = L::vmax(DF[r], Eup);
            Eup = L::vmax(Eup, DFo[r]);
            if constexpr (CAPV) if (r == vrow) *vcap = Eup;
            st.H[r] = best;
            if constexpr (LOCAL) runmax = L::vmax(runmax, best);
            if constexpr (ACC) st.rowacc[r] = L::vmax(st.rowacc[r], best);
            st.F[r] = L::vmax(L::add(st.F[r], vge), L::add(best, vgogo));
        }
    } else {
        // Exact values: the gap opening H + go + ge is shared by both gap states (one add less
        // per cell than the drift-free form of the line above, one op more on the chain).
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const V best = L::vmax(DF[r], Eup);
            const V open = L::add(best, vgogo);
            Eup = L::vmax(L::add(Eup, vge), open);
            if constexpr (CAPV) if (r == vrow) *vcap = Eup;
            st.H[r] = best;
            if constexpr (LOCAL) runmax = L::vmax(runmax, best);
            if constexpr (ACC) st.rowacc[r] = L::vmax(st.rowacc[r], best);
            st.F[r] = L::vmax(L::add(st.F[r], vge), open);
        }
    }
}

// Scores and cells of one step in one go (no read-ahead).
template <int W, int R, int SCORE, bool LOCAL, bool ACC, bool DRIFT, bool CDRIFT>
__device__ __forceinline__ void cell_block(Rows<W, R, ACC> &st, typename Lane<W>::V &Eup,
                                           typename Lane<W>::V diag, typename Lane<W>::V &runmax, uint32_t Bin,
                                           uint32_t tbl_lo, uint32_t tbl_hi, const char *lds, typename Lane<W>::V vgo,
                                           typename Lane<W>::V vgogo, typename Lane<W>::V vge)
{
    ScoreReads<W, R, SCORE> rd;
    issue_scores<W, R, SCORE>(st, Bin, lds, rd);
    typename Lane<W>::V sc[R];
    finish_scores<W, R, SCORE>(st, Bin, tbl_lo, tbl_hi, rd, sc);
    update_cells<W, R, SCORE, LOCAL, ACC, DRIFT, CDRIFT>(st, Eup, diag, runmax, sc, Bin, tbl_lo, tbl_hi, vgo, vgogo,
                                                         vge);
}

// Score capture of one lane in one stripe (global method).  cap: this lane owns row m of half
// q and reads H(m, n) at column n; lcap (free last column): this lane reads the maximum over its
// real rows at column n.  Contributions meet in an int32 atomicMax on the pair's score.  Register
// values drifted by -dge*(r+1) (dge = ge with the row drift, else 0) and by -dcol (the column
// drift of the current step) are shifted back here.
template <int W, int R>
struct Capture {
    int cap[2], capr[2], lcap[2], lvr[2];
    int ev;   // next event column (INT32_MAX: none)

    __device__ __forceinline__ int next(int after) const
    {
        int e = INT32_MAX;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (cap[q] > after && cap[q] < e) e = cap[q];
            if (lcap[q] > after && lcap[q] < e) e = lcap[q];
        }
        return e;
    }

    __device__ __forceinline__ void init(const Task &tk, int row0, int lc)
    {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            cap[q] = -1; capr[q] = 0; lcap[q] = -1; lvr[q] = 0;
            if (q >= Lane<W>::NH || tk.pair[q] == kNone) continue;
            const int m = int(tk.m[q]), n = int(tk.n[q]);
            if (m - 1 >= row0 && m - 1 < row0 + R) { cap[q] = n - 1; capr[q] = m - 1 - row0; }
            if (lc && row0 < m) { lcap[q] = n - 1; lvr[q] = m - row0 < R ? m - row0 : R; }
        }
        ev = next(-1);
    }

    // off0, off1: offsets of the halves' values (int16 re-centring in the wavefront kernel)
    template <bool ACC>
    __device__ __forceinline__ void at_column(int c, const Rows<W, R, ACC> &st, const Task &tk, int32_t *scores,
                                              int dge, int dcol = 0, int off0 = 0, int off1 = 0)
    {
        if (c != ev) return;
#pragma unroll
        for (int q = 0; q < Lane<W>::NH; ++q) {
            const int off = q == 0 ? off0 : off1;
            if (c == cap[q]) {
                int v = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) if (r == capr[q]) v = Lane<W>::half(st.H[r], q);
                atomicMax(scores + tk.pair[q], v + dge * (capr[q] + 1) + dcol + off);
            }
            if (c == lcap[q]) {
                int v = INT32_MIN;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (r < lvr[q]) { const int h = Lane<W>::half(st.H[r], q) + dge * (r + 1) + dcol; v = h > v ? h : v; }
                atomicMax(scores + tk.pair[q], v + off);
            }
        }
        ev = next(c);
    }

    // free last row (the ROWMAX variants only): the owner of row m contributes the maximum of
    // that row
    template <bool ACC>
    __device__ __forceinline__ void row_maximum(const Rows<W, R, ACC> &st, const Task &tk, int32_t *scores,
                                                int dge) const
    {
        if constexpr (ACC) {
#pragma unroll
            for (int q = 0; q < Lane<W>::NH; ++q) {
                if (cap[q] < 0) continue;
                int v = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) if (r == capr[q]) v = Lane<W>::half(st.rowacc[r], q);
                atomicMax(scores + tk.pair[q], v + dge * (capr[q] + 1));
            }
        }
    }
};

// Copies the pair-row table (`entries` dwords) into LDS.
__device__ __forceinline__ void load_lds_pairs(char *lds_raw, const uint32_t *table, uint32_t entries)
{
    uint32_t *lds32 = reinterpret_cast<uint32_t *>(lds_raw);
    for (uint32_t i = threadIdx.x; i < entries; i += blockDim.x) lds32[i] = table[i];
}

// Copies the substitution table into LDS as 32 bank-swizzled replicas: entry e for lane l at
// dword (e >> 1) * 32 + l % 32, half e & 1 (see pa_plan.hpp).
__device__ __forceinline__ void load_lds_table(char *lds_raw, const int16_t *table, uint32_t entries)
{
    int16_t *lds16 = reinterpret_cast<int16_t *>(lds_raw);
    const uint32_t total = entries * 32;
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        const uint32_t e = i >> 5, copy = i & 31;
        lds16[(((e >> 1) << 5) + copy) * 2 + (e & 1)] = table[e];
    }
}

} // namespace pa
