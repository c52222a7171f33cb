// ehd_error.cpp -- the thread-local message behind ehd_last_error(), shared
// by the device-side entry points and the host-only file layer (ehd_io.cpp).
// Host C++ only, so that the file layer builds and runs without HIP.
#include <cstdarg>
#include <cstdio>

#include "../../include/ehd.h"
#include "ehd_error.h"

namespace {
thread_local char g_err[512] = "";
}  // namespace

int ehd::set_error(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char* ehd_last_error(void) { return g_err; }
