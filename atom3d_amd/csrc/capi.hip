// Error reporting and version of the atom3d_amd C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "a3d_common.hpp"

namespace a3d {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace a3d

extern "C" const char* a3d_last_error(void) { return a3d::g_err; }

extern "C" const char* a3d_version(void) { return "atom3d_amd 0.1.0 gfx950"; }
