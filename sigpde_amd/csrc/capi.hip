// Library-wide C ABI pieces: error string and version.
#include "capi_util.h"

namespace sigpde {
char *error_buffer() {
  static thread_local char buf[512] = {0};
  return buf;
}
}  // namespace sigpde

extern "C" const char *sigpde_last_error(void) { return sigpde::error_buffer(); }

extern "C" const char *sigpde_version(void) { return "sigpde_hip 0.1 (gfx950)"; }
