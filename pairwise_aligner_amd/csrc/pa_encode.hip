// pa_encode.hip -- device-side encoding of the residue streams of a plan.
//
// The host planner decides tasks, shapes and offsets from the sequence lengths alone; the bytes
// themselves go to the device once per batch (seq1 bytes, then seq2 bytes) and this kernel turns
// them into the row and column codes the DP kernels read (layout in pa_plan.hpp).  It is byte
// work bound by HBM: every code word is written once, coalesced, and every residue byte is read
// once per task that holds it.
//
// One workgroup per task.  The code tables of the launch (256 row and 256 column codes) sit in
// LDS; pad rows and columns get the pad codes, which score -1 against everything.
#include "pa_kernels.hpp"

namespace pa {

__global__ __launch_bounds__(256) void encode_kernel(EncodeArgs a)
{
    __shared__ uint32_t row_tab[256], col_tab[256];
    row_tab[threadIdx.x] = a.codes->row[threadIdx.x];
    col_tab[threadIdx.x] = a.codes->col[threadIdx.x];
    __syncthreads();

    const uint32_t ti = a.task_begin + blockIdx.x;
    const Task tk = a.tasks[ti];
    const EncodeTask e = a.enc[ti];
    const uint32_t row_pad = a.codes->row_pad, col_pad = a.codes->col_pad;
    const int halves = a.width == 16 ? 2 : 1;
    bool live[2];
    for (int q = 0; q < 2; ++q) live[q] = q < halves && tk.pair[q] != kNone;
    const uint8_t *rsrc[2] = {a.residues + e.row_src[0], a.residues + e.row_src[1]};
    const uint8_t *csrc[2] = {a.residues + e.col_src[0], a.residues + e.col_src[1]};

    // rows: one word per row and lane (the int16 halves hold the two pairs), or for the pair-row
    // table one word per two rows; the query profile needs no row codes
    const uint32_t nrows = tk.S * a.stripe_rows;
    uint32_t *rows = a.rowcodes + tk.row_off;
    if (a.score == kScorePairs) {
        const uint32_t K = a.pair_K;
        for (uint32_t w = threadIdx.x; w < nrows / 2; w += blockDim.x) {
            const uint32_t i = 2 * w;
            uint32_t out = 0;
            for (int q = 0; q < 2; ++q) {
                const uint32_t r1 = live[q] && i < tk.m[q] ? row_tab[rsrc[q][i]] : row_pad;
                const uint32_t r2 = live[q] && i + 1 < tk.m[q] ? row_tab[rsrc[q][i + 1]] : row_pad;
                out |= (((r1 * K + r2) * K * 4u) & 0xFFFFu) << (16 * q);
            }
            rows[w] = out;
        }
    } else if (a.score != kScoreProfile) {
        for (uint32_t i = threadIdx.x; i < nrows; i += blockDim.x) {
            uint32_t out = 0;
            for (int q = 0; q < halves; ++q) {
                const uint32_t code = live[q] && i < tk.m[q] ? row_tab[rsrc[q][i]] : row_pad;
                out |= halves == 2 ? (code & 0xFFFFu) << (16 * q) : code;
            }
            rows[i] = out;
        }
    }

    // columns, padded to whole conveyor chunks plus the prefetch reach of four chunks
    const uint32_t cw = (tk.N + a.G - 1) / a.G * a.G + 4 * a.G;
    uint32_t *cols = a.colcodes + tk.col_off;
    for (uint32_t j = threadIdx.x; j < cw; j += blockDim.x) {
        uint32_t out = 0;
        for (int q = 0; q < halves; ++q) {
            const uint32_t code = live[q] && j < tk.n[q] ? col_tab[csrc[q][j]] : col_pad;
            out |= halves == 2 ? (code & 0xFFFFu) << (16 * q) : code;
        }
        cols[j] = out;
    }
}

hipError_t launch_encode(const EncodeArgs &a, hipStream_t st)
{
    if (a.task_end <= a.task_begin) return hipSuccess;
    hipLaunchKernelGGL(encode_kernel, dim3(a.task_end - a.task_begin), dim3(256), 0, st, a);
    return hipGetLastError();
}

} // namespace pa
