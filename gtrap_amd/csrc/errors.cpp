// Error text storage and version string for the C ABI.
#include <stdarg.h>
#include <stdio.h>
#include "common.h"

namespace gtrap {
static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace gtrap

extern "C" const char *gtrap_last_error(void) { return gtrap::g_err; }
extern "C" const char *gtrap_version(void) { return "gtrap_amd 0.1 (gfx950)"; }
