// The device scratch array of the allocation-free evaluation.  In host code `scratch` only
// names the symbol: its device address (one copy per GPU) comes from hipGetSymbolAddress,
// called once per device with that device current, and only that address is passed to
// kernels.
#include "scratch.hpp"

#include <mutex>

namespace sus2sh {

namespace {

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

}  // namespace

hipError_t scratch_init_current_device()
{
  int device = 0;
  const hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return e;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (r != hipSuccess) return r;
  g_scratch_ptr[device] = p;
  return hipSuccess;
}

void* scratch_device_ptr(int device)
{
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  return g_scratch_ptr[device];
}

}  // namespace sus2sh
