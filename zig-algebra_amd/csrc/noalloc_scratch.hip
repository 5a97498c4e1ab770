// Device scratch of the entry point that allocates nothing
// (zntt_execute_device_noalloc): a file-scope device array, one copy per GPU
// the library runs on, held until the process ends.  Host code must not use
// the array's name as a device address; each device's address is looked up
// once, with that device current, and kept here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"
#include "zntt_gpu.h"

namespace zntt {

namespace {

__device__ __align__(256) uint8_t scratch[ZNTT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` on each device, once looked up

}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_ptr[device] = p;
  return hipSuccess;
}

void *noalloc_scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

}  // namespace zntt
