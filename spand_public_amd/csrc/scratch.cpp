// Static device workspace of the allocation-free preconditioner apply
// (spand_solve_device_noalloc): callers that cannot allocate device memory (code under HIP
// graph capture, serving loops, harnesses that forbid hipMalloc) apply M^{-1} with the
// segments and partial sums placed in this file-scope array instead of the tree's own
// workspace. Every GPU the library runs on holds one copy of SCRATCH_BYTES until the
// process ends; calls that overlap on one GPU share it and race.
#include "scratch.hpp"

#include <mutex>

#include "tree.hpp"

namespace spand {

static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double* g_ptr[MAX_DEVICES] = {};
}  // namespace

// The array's device address on the current device, looked up once per device. In host
// code `scratch` names a host-side stand-in, not device memory: only this address may be
// handed to kernels.
double* scratch_device_ptr() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) throw Error(ERR_ARG, "scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
        g_ptr[dev] = static_cast<double*>(p);
    }
    return g_ptr[dev];
}

}  // namespace spand
