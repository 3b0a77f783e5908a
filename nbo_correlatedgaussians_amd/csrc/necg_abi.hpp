// necg_abi.hpp -- argument checks shared by the C entry points (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <initializer_list>

#include "../../include/necg.h"

namespace necg {

// True when every non-null pointer is memory the GPU can address: device or
// managed allocations, or pinned (registered) host memory.  Plain pageable
// host pointers are refused at the boundary instead of faulting a kernel.
inline bool device_visible(std::initializer_list<const void *> ptrs) {
  for (const void *p : ptrs) {
    if (!p) continue;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    const bool ok = a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
                    (a.type == hipMemoryTypeHost && a.devicePointer != nullptr);
    if (!ok) return false;
  }
  return true;
}

}  // namespace necg

#define NECG_REQUIRE_DEVICE(...)                                      \
  do {                                                                \
    if (!necg::device_visible({__VA_ARGS__})) return NECG_ERR_ARG;    \
  } while (0)
