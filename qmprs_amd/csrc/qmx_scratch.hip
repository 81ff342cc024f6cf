// Library-owned device scratch for the allocation-free sweep entry point
// (qmx_sweep_noalloc).  The array below lives in the code object's zero-
// initialised data: every GPU the library is loaded on holds its own copy
// (kNoAllocScratchBytes) until the process ends, and no call allocates it.
//
// In host code `scratch` names a host-side shadow, not a device address: the
// device address is looked up once per device with hipGetSymbolAddress and
// kept here, and only that address is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qmx_internal.h"

namespace qmx {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of this device's scratch array.  `device` must be the
// calling thread's current device; the first call per device looks the
// symbol up (loading the code object there if needed), later calls return
// the cached address.
hipError_t scratch_address(int device, void** out) {
  if (!out || device < 0 || device >= kMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

}  // namespace qmx
