// ehd_error.h -- the thread-local error message behind ehd_last_error().
#pragma once

namespace ehd {
// Formats the message into the calling thread's error buffer and returns code.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}  // namespace ehd
