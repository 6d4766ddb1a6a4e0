// pa_kernels.hip -- gfx950 kernels of the batched affine-gap (Gotoh) score-only aligner.
//
// Striped wavefront kernel.  One group of G = 16 lanes (one DPP row) sweeps one task: one
// pair in int32 lanes, or two pairs in the two 16-bit halves of every 32-bit lane (packed
// int16, saturating v_pk_add_i16 / v_pk_max_i16).  Lane t of the group owns R consecutive DP
// rows of the current stripe of G*R rows and keeps their H and horizontal-gap state in
// registers.  The lanes move along the DP columns skewed by one step per lane: at step k lane
// t computes column c = k - t for its R rows, receiving from lane t-1 (DPP row_shr:1) the H
// and vertical-gap state of the row above its block and the column residue.  Lane 0 of the
// group receives the stripe's top boundary instead: the initial DP row for the first stripe,
// otherwise the bottom row the previous stripe left in the boundary buffer.  Column residues
// and boundary values reach lane 0 through "conveyors": G consecutive words loaded coalesced,
// one per lane, then shifted down the group by DPP row_shl:1 once per step.
//
// Cell recurrence (rrahn/pairwise_aligner affine/affine_dp_algorithm.hpp:139-156):
//   best = H(i-1,j-1) + s(a_i, b_j)         [local: max(best, 0), dp_matrix_local.hpp:39-47]
//   best = max(best, V(i,j), W(i,j))         H(i,j) = best
//   open = best + go + ge
//   V(i+1,j) = max(V(i,j) + ge, open)        vertical gap state, flows down the rows
//   W(i,j+1) = max(W(i,j) + ge, open)        horizontal gap state, kept per row in registers
// Boundary (affine_initialisation_strategy.hpp:38-51): H(i,0) = first_column free ? 0 : go+ge*i,
// H(0,j) likewise with first_row; the gap state entering the matrix is H + go + ge.
//
// Score extraction (tracker_global_scalar.hpp:36-61, tracker_local_scalar.hpp:38-49):
//   global  H(m,n) is captured by the lane owning row m at the step of column n; with a free
//           last column every lane contributes max over its real rows at column n; with a
//           free last row the per-row maxima are kept and the owner of row m contributes its
//           row's maximum.  Contributions meet in an int32 atomicMax.
//   local   running maximum over all cells, reduced over the group.  Pad rows and columns
//           score -1 against everything, so with gap scores <= 0 no pad cell can exceed the
//           maximum of the real cells.
// Saturation: int16 lanes saturate; a running maximum equal to 32767 flags the pair for an
// int32 recomputation.  Lower bounds are proven on the host before int16 is chosen.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_plan.hpp"
#include "pa_kernels.hpp"

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
};

// DPP controls (gfx9): row_shr:1 moves lane t-1 -> t inside a row of 16 lanes, row_shl:1
// moves lane t+1 -> t.  Lanes without a source keep `old` (bound_ctrl off).
constexpr int kRowShr1 = 0x111;
constexpr int kRowShl1 = 0x101;
constexpr int kRowRor1 = 0x121;
constexpr int kRowRor2 = 0x122;
constexpr int kRowRor4 = 0x124;
constexpr int kRowRor8 = 0x128;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(src), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ int h0_value(int i, int free_flag, int go, int ge)
{
    return (i == 0 || free_flag) ? 0 : go + ge * i;
}

// Smallest capture column greater than `after` (INT32_MAX if none).
__device__ __forceinline__ int next_event(const int cap[2], const int lcap[2], int after)
{
    int e = INT32_MAX;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (cap[q] > after && cap[q] < e) e = cap[q];
        if (lcap[q] > after && lcap[q] < e) e = lcap[q];
    }
    return e;
}

template <int W, int G, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX>
__global__ __launch_bounds__(256) void stripe_kernel(KernelArgs a)
{
    using L = Lane<W>;
    using V = typename L::V;
    extern __shared__ char lds_raw[];

    if constexpr (SCORE == kScoreMatrix) {
        // 32 bank-swizzled copies: entry e for lane l at dword (e >> 1) * 32 + l % 32, half e & 1
        int16_t *lds16 = reinterpret_cast<int16_t *>(lds_raw);
        const uint32_t total = a.table_entries * 32;
        for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
            const uint32_t e = i >> 5, copy = i & 31;
            lds16[(((e >> 1) << 5) + copy) * 2 + (e & 1)] = a.table[e];
        }
        __syncthreads();
    }

    const uint32_t gid = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (gid >= a.ntasks) return;
    const int t = int(threadIdx.x % G);
    const uint32_t lane_bank = (threadIdx.x & 31u) * 4u;   // this lane's copy of the LDS table
    const Task tk = a.tasks[gid];
    const int go = a.go, ge = a.ge;
    // method_local.hpp:41-49: a local alignment has every end gap free
    const int fc = LOCAL ? 1 : a.fc, fr = LOCAL ? 1 : a.fr, lc = LOCAL ? 0 : a.lc;
    const V vgogo = L::splat(go + ge), vge = L::splat(ge);
    const uint32_t tbl_lo = a.tbl_lo, tbl_hi = a.tbl_hi;
    const int N = int(tk.N);
    constexpr int ROWS = G * R;

    V runmax = LOCAL ? L::splat(0) : L::lowest();

    for (uint32_t s = 0; s < tk.S; ++s) {
        const int row0 = int(s) * ROWS + t * R;   // DP row above this lane's block
        uint32_t A[R];
        uint32_t Ahi[R];
        V H[R], F[R];
        V rowacc[ROWMAX ? R : 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t code = a.rowcodes[tk.row_off + uint32_t(row0 + r)];
            if constexpr (SCORE == kScoreMatrix && W == 16) {
                A[r] = (code & 0xFFFFu) + lane_bank;
                Ahi[r] = (code >> 16) + lane_bank;
            } else if constexpr (SCORE == kScoreMatrix) {
                A[r] = code + lane_bank;
                Ahi[r] = 0;
            } else {
                A[r] = code;
                Ahi[r] = 0;
            }
            H[r] = L::splat(h0_value(row0 + r + 1, fc, go, ge));
            F[r] = L::add(H[r], vgogo);
            if constexpr (ROWMAX) rowacc[r] = L::lowest();
        }
        V Hdiag = L::splat(h0_value(row0, fc, go, ge));

        // Capture events of the two halves (global), as column indices (-1: none).  cap: this
        // lane owns row m and reads H(m, n) at column n; lcap (free last column): this lane
        // reads the maximum over its real rows at column n.  `ev` is the next event column.
        int cap[2] = {-1, -1}, capr[2] = {0, 0};
        int lcap[2] = {-1, -1}, lvr[2] = {0, 0};
#pragma unroll
        for (int q = 0; q < L::NH; ++q) {
            if (tk.pair[q] == kNone) continue;
            const int m = int(tk.m[q]), n = int(tk.n[q]);
            if (m - 1 >= row0 && m - 1 < row0 + R) { cap[q] = n - 1; capr[q] = m - 1 - row0; }
            if (lc && row0 < m) { lcap[q] = n - 1; lvr[q] = m - row0 < R ? m - row0 : R; }
        }
        int ev = next_event(cap, lcap, -1);

        const bool has_bnd_in = s > 0;
        const bool has_bnd_out = s + 1 < tk.S;
        const uint32_t *colp = a.colcodes + tk.col_off;
        uint32_t *bndp = has_bnd_in || has_bnd_out ? a.bnd + 2ull * tk.bnd_off : nullptr;

        // conveyors: current chunk and the next one
        uint32_t convB = colp[t], nextB = colp[G + t];
        uint32_t convH = 0, convE = 0, nextH = 0, nextE = 0;
        if (has_bnd_in) {
            convH = bndp[2 * t]; convE = bndp[2 * t + 1];
            nextH = bndp[2 * (G + t)]; nextE = bndp[2 * (G + t) + 1];
        }
        // outputs of this lane from the previous step (read by lane t+1)
        uint32_t Hout = 0, Eout = 0, Bout = 0;
        int hb_top = h0_value(1, fr, go, ge);   // H(0, k+1) for the first stripe

        const int nsteps = N + G - 1;
        int c = -t;   // DP column (0-based) this lane computes at step k
        for (int k = 0; k < nsteps; ++k, ++c) {
            if (k > 0 && (k % G) == 0) {
                convB = nextB;
                nextB = colp[k + G + t];
                if (has_bnd_in) {
                    convH = nextH; convE = nextE;
                    nextH = bndp[2 * (k + G + t)];
                    nextE = bndp[2 * (k + G + t) + 1];
                }
            }
            uint32_t topH, topE;
            if (has_bnd_in) {
                topH = convH;
                topE = convE;
            } else {
                // H(0, k+1) and the vertical gap state below it: uniform, computed in scalar
                // registers.  Both are <= 0 because go + ge <= 0 and ge <= 0.
                const int e = hb_top + go + ge;
                if constexpr (W == 16) {
                    topH = (uint32_t(hb_top < -32768 ? -32768 : hb_top) & 0xFFFFu) * 0x10001u;
                    topE = (uint32_t(e < -32768 ? -32768 : e) & 0xFFFFu) * 0x10001u;
                } else {
                    topH = uint32_t(hb_top);
                    topE = uint32_t(e);
                }
                hb_top = fr ? 0 : hb_top + ge;
            }
            const uint32_t Bin = dpp<kRowShr1>(convB, Bout);
            const uint32_t Hin = dpp<kRowShr1>(topH, Hout);
            const uint32_t Ein = dpp<kRowShr1>(topE, Eout);
            convB = dpp<kRowShl1>(convB, convB);
            if (has_bnd_in) {
                convH = dpp<kRowShl1>(convH, convH);
                convE = dpp<kRowShl1>(convE, convE);
            }
            if (uint32_t(c) < uint32_t(N)) {
                V Eup = L::from(Ein);
                V diag = Hdiag;
                uint32_t Blo = Bin, Bhi = 0;
                if constexpr (SCORE == kScoreMatrix && W == 16) { Blo = Bin & 0xFFFFu; Bhi = Bin >> 16; }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    V sc;
                    if constexpr (SCORE == kScoreDna) {
                        sc = L::from(__builtin_amdgcn_perm(tbl_hi, tbl_lo, A[r] ^ Bin));
                    } else if constexpr (W == 16) {
                        const short s0 = *reinterpret_cast<const short *>(lds_raw + A[r] + Blo);
                        const short s1 = *reinterpret_cast<const short *>(lds_raw + Ahi[r] + Bhi);
                        sc = V{s0, s1};
                    } else {
                        sc = L::splat(int(*reinterpret_cast<const short *>(lds_raw + A[r] + Blo)));
                    }
                    V best = L::add(diag, sc);
                    if constexpr (LOCAL) best = L::vmax(best, L::splat(0));
                    best = L::vmax(L::vmax(best, Eup), F[r]);
                    diag = H[r];
                    H[r] = best;
                    if constexpr (LOCAL || TRACK) runmax = L::vmax(runmax, best);
                    if constexpr (ROWMAX) rowacc[r] = L::vmax(rowacc[r], best);
                    const V open = L::add(best, vgogo);
                    Eup = L::vmax(L::add(Eup, vge), open);
                    F[r] = L::vmax(L::add(F[r], vge), open);
                }
                Hdiag = L::from(Hin);
                Hout = L::bits(H[R - 1]);
                Eout = L::bits(Eup);
                Bout = Bin;
                if (has_bnd_out) {
                    if (t == G - 1) {
                        bndp[2 * c] = Hout;
                        bndp[2 * c + 1] = Eout;
                    }
                }
                if constexpr (!LOCAL) {
                    if (c == ev) {
#pragma unroll
                        for (int q = 0; q < L::NH; ++q) {
                            if (c == cap[q]) {
                                int v = 0;
#pragma unroll
                                for (int r = 0; r < R; ++r) if (r == capr[q]) v = L::half(H[r], q);
                                atomicMax(a.scores + tk.pair[q], v);
                            }
                            if (c == lcap[q]) {
                                int v = INT32_MIN;
#pragma unroll
                                for (int r = 0; r < R; ++r)
                                    if (r < lvr[q]) { const int h = L::half(H[r], q); v = h > v ? h : v; }
                                atomicMax(a.scores + tk.pair[q], v);
                            }
                        }
                        ev = next_event(cap, lcap, c);
                    }
                }
            }
        }
        if constexpr (ROWMAX) {
#pragma unroll
            for (int q = 0; q < L::NH; ++q) {
                if (cap[q] < 0) continue;
                int v = 0;
#pragma unroll
                for (int r = 0; r < R; ++r) if (r == capr[q]) v = L::half(rowacc[r], q);
                atomicMax(a.scores + tk.pair[q], v);
            }
        }
        if (has_bnd_out) {
            // the boundary row written by lane G-1 is read by lane 0 of the next stripe
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }

    if constexpr (LOCAL || TRACK) {
        // maximum over the 16 lanes of the group
        uint32_t m = L::bits(runmax);
        m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor8>(m, m))));
        m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor4>(m, m))));
        m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor2>(m, m))));
        m = L::bits(L::vmax(L::from(m), L::from(dpp<kRowRor1>(m, m))));
        if (t == 0) {
            const V mv = L::from(m);
#pragma unroll
            for (int q = 0; q < L::NH; ++q) {
                if (tk.pair[q] == kNone) continue;
                const int v = L::half(mv, q);
                if constexpr (LOCAL) a.scores[tk.pair[q]] = v;
                if (W == 16 && v >= L::kMax) atomicOr(a.flags + tk.pair[q], kFlagSaturated);
            }
        }
    }
}

// Scores of pairs with an empty sequence (closed forms) and the initial value of every other
// score: INT32_MIN, raised by the end-gap candidates H(m,0) (free last row) and H(0,n) (free
// last column) that no kernel visits.
__global__ void init_scores_kernel(InitArgs a)
{
    const uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n_pairs) return;
    const uint64_t p = a.subset ? a.subset[idx] : idx;
    const PairMeta pm = a.pairs[p];
    const int64_t go = a.go, ge = a.ge;
    const int64_t m = pm.m, n = pm.n;
    uint32_t flags = (pm.width == 16 ? 0x01u : (pm.width == 32 ? 0x02u : 0x08u)) | a.flag_or;
    int64_t s = INT32_MIN;
    if (pm.local) {
        s = INT32_MIN;
    } else {
        const int fc = a.fc, fr = a.fr, lc = a.lc, lr = a.lr;
        auto h0row = [&](int64_t j) -> int64_t { return (j == 0 || fr) ? 0 : go + ge * j; };  // H(0,j)
        auto h0col = [&](int64_t i) -> int64_t { return (i == 0 || fc) ? 0 : go + ge * i; };  // H(i,0)
        if (m == 0) {
            s = h0row(n);
            if (lr && n > 0) { s = s > 0 ? s : 0; s = s > h0row(1) ? s : h0row(1); }
        } else if (n == 0) {
            s = h0col(m);
            if (lc) { s = s > 0 ? s : 0; s = s > h0col(1) ? s : h0col(1); }
        } else {
            if (lr) s = s > h0col(m) ? s : h0col(m);
            if (lc) s = s > h0row(n) ? s : h0row(n);
        }
    }
    a.scores[p] = int32_t(s);
    a.flags[p] = flags;
}

template <int W, int G, int R, int SCORE, bool LOCAL, bool TRACK, bool ROWMAX>
static hipError_t launch_one(const KernelArgs &a, uint32_t lds_bytes, hipStream_t st)
{
    const uint32_t threads = 256;
    const uint64_t total = uint64_t(a.ntasks) * G;
    const uint32_t blocks = uint32_t((total + threads - 1) / threads);
    hipLaunchKernelGGL((stripe_kernel<W, G, R, SCORE, LOCAL, TRACK, ROWMAX>), dim3(blocks), dim3(threads),
                       SCORE == kScoreMatrix ? lds_bytes : 0, st, a);
    return hipGetLastError();
}

template <int W, int R, int SCORE, bool LOCAL>
static hipError_t dispatch_flags(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    if (LOCAL) return launch_one<W, 16, R, SCORE, LOCAL, false, false>(a, lds, st);
    if (k.rowmax) {
        if (k.track) return launch_one<W, 16, R, SCORE, LOCAL, true, true>(a, lds, st);
        return launch_one<W, 16, R, SCORE, LOCAL, false, true>(a, lds, st);
    }
    if (k.track) return launch_one<W, 16, R, SCORE, LOCAL, true, false>(a, lds, st);
    return launch_one<W, 16, R, SCORE, LOCAL, false, false>(a, lds, st);
}

template <int W, int R>
static hipError_t dispatch_score(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    if (k.score == kScoreDna) {
        return k.local ? dispatch_flags<W, R, kScoreDna, true>(k, a, lds, st)
                       : dispatch_flags<W, R, kScoreDna, false>(k, a, lds, st);
    }
    return k.local ? dispatch_flags<W, R, kScoreMatrix, true>(k, a, lds, st)
                   : dispatch_flags<W, R, kScoreMatrix, false>(k, a, lds, st);
}

template <int W>
static hipError_t dispatch_rows(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    switch (k.R) {
    case 4: return dispatch_score<W, 4>(k, a, lds, st);
    case 6: return dispatch_score<W, 6>(k, a, lds, st);
    case 8: return dispatch_score<W, 8>(k, a, lds, st);
    case 10: return dispatch_score<W, 10>(k, a, lds, st);
    case 12: return dispatch_score<W, 12>(k, a, lds, st);
    case 14: return dispatch_score<W, 14>(k, a, lds, st);
    case 16: return dispatch_score<W, 16>(k, a, lds, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stripe(const VariantKey &k, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st)
{
    if (k.G != 16 || a.ntasks == 0) return k.G != 16 ? hipErrorInvalidValue : hipSuccess;
    return k.width == 16 ? dispatch_rows<16>(k, a, lds_bytes, st) : dispatch_rows<32>(k, a, lds_bytes, st);
}

hipError_t launch_init(const InitArgs &a, hipStream_t st)
{
    if (a.n_pairs == 0) return hipSuccess;
    const uint32_t threads = 256;
    const uint32_t blocks = uint32_t((a.n_pairs + threads - 1) / threads);
    hipLaunchKernelGGL(init_scores_kernel, dim3(blocks), dim3(threads), 0, st, a);
    return hipGetLastError();
}

} // namespace pa
