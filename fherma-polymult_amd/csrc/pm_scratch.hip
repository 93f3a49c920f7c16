// Workspace of the allocation-free product entry point (pm_multiply_device_noalloc):
// the two K x N residue arrays of one product live in a file-scope device array,
// which the code object's loader places on every GPU the library runs on and
// which stays there until the process ends. No call on this path allocates.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device with hipGetSymbolAddress (at plan
// creation, and on a device's first call if no plan was created there) and only
// that address is ever handed to kernels.
#include "pm_state.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

namespace pm {
namespace {

static __device__ __align__(256) uint32_t scratch[kScratchWords];

std::mutex g_scratch_mutex;
std::vector<uint32_t*> g_scratch_ptr;   // per device id, null until looked up

}  // namespace

uint32_t* scratch_on_current_device() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (static_cast<size_t>(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw HipError("hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace pm
