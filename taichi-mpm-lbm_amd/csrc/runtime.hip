// Error reporting and version entry points of the C ABI.
#include <cstdarg>

#include "common.h"

namespace fsi {

namespace {
thread_local char g_error[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_error; }

int check_launch(const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", where, hipGetErrorString(e));
    return FSI_ERR_HIP;
  }
  return 0;
}

}  // namespace fsi

extern "C" {

int fsi_abi_version(void) { return FSI_ABI_VERSION; }

const char* fsi_last_error(void) { return fsi::last_error(); }

}  // extern "C"
