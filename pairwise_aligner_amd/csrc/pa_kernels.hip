// pa_kernels.hip -- launches of the gfx950 kernels: the initial-score kernel and the routing of a
// DP launch to the instantiation of its lane width and rows per lane (pa_kernels_inst.hip; the
// kernels themselves are in pa_dp_kernels.hpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_kernels.hpp"
#include "pa_plan.hpp"

namespace pa {

// Defined once per (W, R) in pa_kernels_inst.hip.
template <int W, int R>
hipError_t dispatch_score(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st);

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
    if (!pm.local) {
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
    if (idx == 0 && a.subset == nullptr) a.flags[a.n_pairs] = 0;   // error slot of the wavefront kernel
}

template <int W>
static hipError_t dispatch_rows(const VariantKey &k, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    if (k.G == 64) {   // the wavefront kernel runs stripes of 64 x 16 rows only
        if (k.R == 16) return dispatch_score<W, 16>(k, a, lds, st);
        return hipErrorInvalidValue;
    }
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
    if ((k.G != 16 && k.G != 64) || a.ntasks == 0) return a.ntasks == 0 ? hipSuccess : hipErrorInvalidValue;
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
