// pa_fused.hip -- launches of the fused striped kernel (pa_dp_kernels.hpp, fused_stripe_kernel):
// int16 lanes; the local method with the byte-permute DNA scores or the pair-row table, every
// rows-per-lane R in one kernel; the global method with the byte-permute scores, R <= 12 and
// R >= 14 in two kernels (pa_plan.hpp, fused_launchable, fused_rows).
#include "pa_dp_kernels.hpp"

namespace pa {

template <typename K>
static hipError_t launch_one(K kernel, const KernelArgs &a, uint32_t lds, hipStream_t st)
{
    const uint32_t blocks = (a.nitems + 3) / 4;   // four waves per block, one item per wave at least
    hipLaunchKernelGGL(kernel, striped_grid(kernel, a, blocks, lds), dim3(256), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_fused(const VariantKey &k, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st)
{
    if (a.nitems == 0) return hipSuccess;
    if (!a.queue || !a.items || !fused_launchable(k)) return hipErrorInvalidValue;
    if (k.local) {
        if (k.score == kScoreDna) return launch_one(fused_stripe_kernel<16, kScoreDna, true, 4, 16>, a, 0, st);
        return launch_one(fused_stripe_kernel<16, kScorePairs, true, 4, 16>, a, lds_bytes, st);
    }
    if (fused_rows(k).lo == 4) return launch_one(fused_stripe_kernel<16, kScoreDna, false, 4, 12>, a, 0, st);
    return launch_one(fused_stripe_kernel<16, kScoreDna, false, 14, 16>, a, 0, st);
}

} // namespace pa
