// Error reporting and library-level entry points of the C ABI.
#include "ua_common.h"

#include <string.h>

namespace ua {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_error; }

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return UA_ERR_HIP;
  }
  return UA_OK;
}

}  // namespace ua

extern "C" int ua_abi_version(void) { return UA_ABI_VERSION; }

extern "C" const char* ua_last_error(void) { return ua::get_error(); }
