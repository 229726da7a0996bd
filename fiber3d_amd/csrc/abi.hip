// Error reporting and version query of the libf3d C ABI.
#include <stdarg.h>

#include "common.h"

namespace f3d {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

}  // namespace f3d

extern "C" const char* f3d_last_error(void) { return f3d::g_last_error; }

extern "C" int f3d_abi_version(void) { return 100; }
