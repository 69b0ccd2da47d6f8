// The fixed device scratch of the no-allocation entry point (tb_syrk_noalloc).
//
// Some callers cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  For them the library keeps one
// file-scope device array of kNoAllocScratchBytes; every GPU that loads the library holds
// its own copy until the process ends.  In host code the array's name is only a host-side
// stand-in: its device address is read once per device with hipGetSymbolAddress (with that
// device current, from tb_create) and only that address is ever handed to kernels.
//
// Every call on a GPU shares that GPU's array and leaves its contents behind: calls that
// overlap on one GPU (two streams, two threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tb_kernels.hpp"

namespace {

static __device__ __align__(256) char scratch[tb::kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

namespace tb {

hipError_t noalloc_scratch(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace tb
