// Error reporting and library identification for the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "common.h"

namespace matpl {
namespace {
thread_local char g_error[512] = "";
}

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}
}  // namespace matpl

MATPL_EXPORT const char *matpl_last_error(void) { return matpl::g_error; }

MATPL_EXPORT int matpl_abi_version(void) { return 2; }
