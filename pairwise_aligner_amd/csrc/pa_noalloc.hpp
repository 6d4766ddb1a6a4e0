// pa_noalloc.hpp -- the per-device scratch of pa_score_batch_noalloc (defined in pa_noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "pa_device.h"

namespace pa {

inline constexpr size_t kNoallocScratchBytes = PA_NOALLOC_SCRATCH_BYTES;

// With `device` current: looks up the address of that device's copy of the scratch array
// (hipGetSymbolAddress) once and keeps it.  Called by pa_ctx_create.
hipError_t noalloc_setup(int device);
// The address noalloc_setup found for `device`, or nullptr.
void *noalloc_scratch(int device);
// Held for the whole of a pa_score_batch_noalloc call on `device`.
std::mutex &noalloc_call_mutex(int device);

} // namespace pa
