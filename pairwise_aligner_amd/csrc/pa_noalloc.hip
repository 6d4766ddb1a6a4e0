// pa_noalloc.hip -- the device memory of pa_score_batch_noalloc (include/pa_device.h): one
// static array per device, in place of every hipMalloc of a one-shot call (residues, tasks,
// code streams, substitution tables, boundary rows, scores, flags and the int32 re-plan of
// saturated pairs, bump-allocated in that order by pa_capi.cpp).
//
// In host code `scratch` names a host-side stand-in, not device memory: host code uses only the
// address hipGetSymbolAddress returns for the current device, kept per device below.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pa_noalloc.hpp"

namespace pa {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_setup;
void *g_scratch[kMaxDevices] = {};
std::mutex g_call[kMaxDevices];
} // namespace

hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup);
    if (g_scratch[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch[device] = ptr;
    return hipSuccess;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_setup);
    return g_scratch[device];
}

std::mutex &noalloc_call_mutex(int device) { return g_call[device]; }

} // namespace pa
