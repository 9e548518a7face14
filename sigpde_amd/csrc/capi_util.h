// Error reporting shared by the C ABI entry points: a thread-local message plus status codes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/sigpde_hip.h"

namespace sigpde {

// Refined rows (m << order) and columns (n << order) of one pair. Rows cost nothing on chip;
// columns are held by the band-boundary lines (LDS, or global memory past what LDS holds). Both
// are bounded so that a wave's virtual-column count (bands x band length) stays within int.
constexpr int64_t kMaxRefinedRows = 1 << 20;
constexpr int64_t kMaxRefinedCols = 1 << 22;
constexpr int64_t kMaxVirtualColumns = (int64_t)1 << 30;
// LDS of one workgroup on gfx950 (160 KB).
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

// Workgroups of `kernel` (threads per workgroup, lds bytes of LDS) one CU of the current device
// holds at once (0 when the query fails).
inline int blocks_per_cu(const void *kernel, int threads, size_t lds) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess)
    return 0;
  return per;
}

// CUs of the current device (0 when the query fails).
inline int cu_count() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return cus;
}

inline int hip_status(hipError_t e, const char *what) {
  if (e == hipSuccess) return SIGPDE_OK;
  return set_error(e == hipErrorOutOfMemory ? SIGPDE_ENOMEM : SIGPDE_EDEVICE, "%s: %s", what,
                   hipGetErrorString(e));
}

}  // namespace sigpde
