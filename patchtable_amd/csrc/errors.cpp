// Thread-local last-error string of the C ABI (replaces ASSERT2's fprintf + exit(1),
// solver/util.h:33).
#include "pt_internal.h"

#include <cstdarg>
#include <cstdio>

namespace {
thread_local char g_last_error[1024] = "";
}

int pt_fail(int code, const char *msg) {
    snprintf(g_last_error, sizeof(g_last_error), "%s", msg ? msg : "");
    return code;
}

int pt_failf(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *pt_last_error(void) { return g_last_error; }
