// pa_fused.hip -- launches of the fused striped kernel (pa_dp_kernels.hpp, fused_stripe_kernel):
// int16 lanes, local method, byte-permute DNA scores or the pair-row table (pa_plan.hpp,
// fused_launchable).  The global method is not fused: its variants at R = 16 hold 232 VGPRs
// alone, and the kernel that holds all R would run one wave per SIMD.
#include "pa_dp_kernels.hpp"

namespace pa {

hipError_t launch_fused(const VariantKey &k, const KernelArgs &a, uint32_t lds_bytes, hipStream_t st)
{
    if (a.nitems == 0) return hipSuccess;
    if (!a.queue || !a.items || !fused_launchable(k)) return hipErrorInvalidValue;
    const uint32_t blocks = (a.nitems + 3) / 4;   // four waves per block, one item per wave at least
    if (k.score == kScoreDna) {
        auto kernel = fused_stripe_kernel<16, kScoreDna, true>;
        hipLaunchKernelGGL(kernel, striped_grid(kernel, a, blocks, 0), dim3(256), 0, st, a);
    } else {
        auto kernel = fused_stripe_kernel<16, kScorePairs, true>;
        hipLaunchKernelGGL(kernel, striped_grid(kernel, a, blocks, lds_bytes), dim3(256), lds_bytes, st, a);
    }
    return hipGetLastError();
}

} // namespace pa
