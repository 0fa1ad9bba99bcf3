// Sweep state of the allocation-free entry point (bcmi_bc_run_static): one
// static device array per GPU.  It sits in a translation unit, and so a code
// object, of its own: a process loads its 64 MiB only when it first calls that
// entry point, and every other path keeps the engine's small code object.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"

namespace bcmi {

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

static std::mutex g_scratch_init_mu;
static void* g_scratch_ptr[kMaxDevices];  // this device's copy of `scratch` (hipGetSymbolAddress)

int static_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return set_error(BCMI_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_init_mu);
  if (g_scratch_ptr[device]) return BCMI_OK;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess || cur != device) return set_error(BCMI_ERR_STATE, "static scratch set up with another device current");
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(BCMI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return BCMI_OK;
}

void* static_scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_init_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

}  // namespace bcmi
