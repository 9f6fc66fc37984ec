// rfa_noalloc.hip -- the static device workspace of rfa_pairs_noalloc.
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy from module load to process end.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol); it
// takes the device address once per device with hipGetSymbolAddress, with that
// device current, and uses that address for every launch.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rfa_internal.hpp"

namespace rfa {

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch_address(int device, void** ptr)
{
    static std::mutex mu;
    static void* addr[NOALLOC_MAX_DEVICES] = {};
    if (device < 0 || device >= (int)NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)); // the caller made `device` current
        if (e != hipSuccess) return e;
        addr[device] = p;
    }
    *ptr = addr[device];
    return hipSuccess;
}

} // namespace rfa
