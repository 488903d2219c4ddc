// noalloc.hip -- the fixed workspace of the allocation-free TSMQR
// (avs_tsmqr_noalloc).
//
// `scratch` holds the T_JJ factor (16 * b doubles) that avs_tsmqr builds in a
// hipMalloc'd buffer.  Each GPU the library runs on holds its own copy until
// the process ends.  In host code the name `scratch` is only a handle on the
// symbol, not a device address: noalloc_scratch() looks the address up with
// hipGetSymbolAddress once per device, with that device current, and every
// kernel argument uses the address it returns.
#include <hip/hip_runtime.h>

#include <mutex>
#include <unordered_map>

#include "kernels.hpp"

namespace avs {

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

hipError_t noalloc_scratch(int device, double **ptr)
{
    static std::mutex mu;
    static std::unordered_map<int, double *> addr; // device ordinal -> its copy of `scratch`
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) {
        const hipError_t r = hipSetDevice(cur);
        if (e == hipSuccess) e = r;
    }
    if (e != hipSuccess) return e;
    addr[device] = static_cast<double *>(p);
    *ptr = static_cast<double *>(p);
    return hipSuccess;
}

} // namespace avs
