// Static device workspace of the allocation-free apply (scratch.cpp).
#pragma once
#include <cstddef>

namespace spand {

// Size of the static workspace: segments (N x nrhs doubles) plus the mv phases' partial
// sums must fit, e.g. N = 262,144 with up to 30 right-hand sides.
constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;

// Device address of the workspace on the current device (looked up on first use per device).
double* scratch_device_ptr();

}  // namespace spand
