// rkdg_noalloc.hip — the static device scratch array behind rkdg_step_noalloc.
//
// The array is a file-scope __device__ object: every GPU the library runs on
// holds its own copy, from module load until the process ends. Host code never
// passes the array's name to a kernel (on the host that name is only a
// stand-in, not a device address): noalloc_scratch() looks up the address on
// the current device with hipGetSymbolAddress, once per device, and the
// entry point uses that pointer for every kernel argument.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rkdg_kernels.hpp"

namespace rkdg {

// 16 Mi doubles = 128 MiB: four stage vectors of up to 4 Mi doubles each
// (for example a 24^3 Q2 hex mesh, or 2D meshes of about 250k Q1 elements)
static __device__ __align__(256) double scratch[kScratchDoubles];

hipError_t noalloc_scratch(double** out) {
  static std::mutex mu;
  static double* ptr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[dev] = static_cast<double*>(p);
  }
  *out = ptr[dev];
  return hipSuccess;
}

}  // namespace rkdg
