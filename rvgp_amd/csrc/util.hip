// Error reporting and version for the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace rvgp {
static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace rvgp

extern "C" int rvgp_version(void) { return 1; }

extern "C" const char* rvgp_last_error(void) { return rvgp::g_err; }
