// Fixed device scratch for the allocation-free evaluation (sus2sh_compute_noalloc): one
// file-scope device array per GPU, held until the process ends, whose address is looked up
// once per device when a handle is created on it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace sus2sh {

// 128 MiB: the work buffers of one evaluation of ~65,000 atoms of an l3k3 (rb 8, rc 5.0)
// model in FP32, or ~30,000 in FP64.
constexpr size_t kScratchBytes = size_t(128) << 20;

// Look up the scratch array's device address on the current device (once per device; later
// calls return hipSuccess without a runtime call).
hipError_t scratch_init_current_device();

// Device address of the scratch array on `device`, or nullptr before
// scratch_init_current_device() ran with that device current.
void* scratch_device_ptr(int device);

}  // namespace sus2sh
