// pp_memcheck.hpp -- the Tier B memory contract, checked on the host.
//
// Tier B arrays must be device memory of the current device allocated with hipMalloc
// (coarse-grained; what torch, hipMalloc and hipMallocAsync hand out). The scatters end
// in device-scope float atomics (global_atomic_add_f32), which are exact read-modify-
// writes only on coarse-grained device memory; on host, managed or unified memory they
// may be lost or split. Fine-grained device memory (hipExtMallocWithFlags with
// hipDeviceMallocFinegrained) is outside the contract as well but cannot be told apart
// here from its pointer attributes.
#pragma once
#include <hip/hip_runtime.h>

namespace pp {

// nullptr when p lies in device memory of device `dev`; else a short reason.
inline const char* device_memory_issue(const void* p, int dev) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();  // clear the lookup's error so the next call does not report it
    return "not memory allocated through HIP";
  }
  if (at.isManaged || at.type == hipMemoryTypeManaged || at.type == hipMemoryTypeUnified)
    return "managed or unified memory";
  if (at.type == hipMemoryTypeHost) return "host memory";
  if (at.type != hipMemoryTypeDevice) return "not device memory";
  if (at.device != dev) return "memory of another device";
  return nullptr;
}

}  // namespace pp
