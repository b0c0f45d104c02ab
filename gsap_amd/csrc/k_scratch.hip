// k_scratch.hip -- the static device scratch of gsap_sap_collide_noalloc.
//
// Every update buffer of one allocation-free call is carved from `scratch`.
// Each GPU the library runs on holds its own copy, GSAP_NOALLOC_SCRATCH_BYTES
// long, until the process ends.  Host code never passes the array's name to
// a kernel or a copy (in host code it names a host-side stand-in): it uses the
// device address hipGetSymbolAddress returns, looked up once per device with
// that device current and kept in g_ptr.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gsap.h"
#include "gsap_internal.h"

namespace gsap {

static __device__ __attribute__((aligned(256))) unsigned char scratch[GSAP_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_region(void** ptr, size_t* bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p || ((uintptr_t)p & 255u)) return hipErrorInvalidSymbol;
        g_ptr[dev] = p;
    }
    *ptr = g_ptr[dev];
    *bytes = (size_t)GSAP_NOALLOC_SCRATCH_BYTES;
    return hipSuccess;
}

}  // namespace gsap
