// Library-wide ABI glue: version and the per-thread error message.
#include <stdarg.h>
#include <stdio.h>

#include "dfcsr.h"

namespace dfcsr {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace dfcsr

extern "C" int dfcsr_abi_version(void) { return DFCSR_ABI_VERSION; }

extern "C" const char *dfcsr_last_error(void) { return dfcsr::g_err; }
