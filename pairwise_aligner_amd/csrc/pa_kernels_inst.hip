// pa_kernels_inst.hip -- one instantiation of the DP kernels (pa_dp_kernels.hpp): lane width
// PA_INST_W and rows per lane PA_INST_R, both given on the compiler command line.  The build
// compiles this file once per (W, R) pair in parallel.
#include "pa_dp_kernels.hpp"

#if !defined(PA_INST_W) || !defined(PA_INST_R)
#error "compile with -DPA_INST_W=<16|32> -DPA_INST_R=<4..16>"
#endif

namespace pa {
template hipError_t dispatch_score<PA_INST_W, PA_INST_R>(const VariantKey &, const KernelArgs &, uint32_t,
                                                         hipStream_t);
} // namespace pa
