// Error reporting shared by the C ABI entry points: a thread-local message plus status codes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/sigpde_hip.h"

namespace sigpde {

// Refined rows (m << order) of one pair: bounded only so that the kernels' virtual-column
// counts (bands x band length x pairs per stream) stay within int.
constexpr int64_t kMaxRefinedRows = 1 << 20;
// LDS of one workgroup on gfx950 (160 KB). It bounds the refined columns (n << order) of a pair:
// the wave keeps its band-boundary lines, one value per refined column, in LDS.
constexpr size_t kMaxLdsBytes = 163840;

char *error_buffer();  // thread-local, defined in capi.hip

inline void clear_error() { error_buffer()[0] = '\0'; }

__attribute__((format(printf, 2, 3))) inline int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(error_buffer(), 512, fmt, ap);
  va_end(ap);
  return code;
}

inline int hip_status(hipError_t e, const char *what) {
  if (e == hipSuccess) return SIGPDE_OK;
  return set_error(e == hipErrorOutOfMemory ? SIGPDE_ENOMEM : SIGPDE_EDEVICE, "%s: %s", what,
                   hipGetErrorString(e));
}

}  // namespace sigpde
