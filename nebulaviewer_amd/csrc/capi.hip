// capi.hip - error plumbing of the C ABI (thread-local last-error message) and the ABI version.
#include <stdarg.h>
#include "nv_common.h"

namespace nv {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return NV_ERR_LAUNCH;
  }
  return NV_OK;
}

}  // namespace nv

extern "C" const char* nv_last_error(void) { return nv::g_err; }
extern "C" int nv_abi_version(void) { return 5; }
