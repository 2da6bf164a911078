// vc_abi.hip -- error reporting and version of the C ABI (include/vc_abi.h).
#include "vc_common.h"
#include <stdarg.h>
#include <stdio.h>

static thread_local char g_vc_err[512] = "";

void vc_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_vc_err, sizeof(g_vc_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *vc_last_error(void) { return g_vc_err; }

extern "C" int vc_version(void) { return 1; }
