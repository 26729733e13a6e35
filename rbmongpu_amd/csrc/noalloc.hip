// Device scratch of the allocation-free entry point (rbmhip_local_energy_noalloc).
//
// The array lives in the library's code object: every GPU the library runs on
// holds its own copy, of NOALLOC_SCRATCH_BYTES, from the first use of the
// library on that device until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in symbol); the
// per-device address comes from hipGetSymbolAddress, called once per device
// with that device current, when the library sets the device up.
#include <hip/hip_runtime.h>

#include "noalloc.hpp"

namespace rbmhip {

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

hipError_t noalloc_scratch_address(void** ptr) {
    *ptr = nullptr;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    *ptr = p;
    return hipSuccess;
}

}  // namespace rbmhip
