// Device address of this device's fixed scratch array (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace jq {
hipError_t noalloc_scratch(void** out);
}  // namespace jq
