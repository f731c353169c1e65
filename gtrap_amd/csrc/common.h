// Shared host/device helpers for libgtrap_hip.so (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include "../../include/gtrap_hip.h"

namespace gtrap {

// Per-thread last-error text (gtrap_last_error).
void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Map a HIP launch status to a library status, recording the text.
inline int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return GTRAP_ELAUNCH;
  }
  return GTRAP_OK;
}

#define GTRAP_REQUIRE(cond, ...)                                  \
  do {                                                            \
    if (!(cond)) {                                                \
      ::gtrap::set_error(__VA_ARGS__);                            \
      return GTRAP_EINVAL;                                        \
    }                                                             \
  } while (0)

constexpr int kWave = 64;

__device__ __forceinline__ float med3(float a, float b, float c) {
  return __builtin_amdgcn_fmed3f(a, b, c);
}

}  // namespace gtrap
