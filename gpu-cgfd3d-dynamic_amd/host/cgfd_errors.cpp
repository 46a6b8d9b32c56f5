/* cgfd_errors.cpp -- per-thread last-error state for the C ABI. */
#include "cgfd_errors.h"
#include "../../include/cgfd_core.h"

#include <cstdarg>
#include <cstdio>

namespace {
thread_local char g_msg[1024] = "";
}

int cgfd_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  return code;
}

void cgfd_clear_error(void) { g_msg[0] = '\0'; }

extern "C" const char *cgfd_last_error(void) { return g_msg; }

extern "C" const char *cgfd_version(void) { return "cgfd-mi355x 0.1.0 (gfx950)"; }
