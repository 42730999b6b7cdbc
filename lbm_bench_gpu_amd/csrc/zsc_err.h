// zsc_err.h -- per-thread error message of the C-ABI (internal).
#pragma once

namespace zsc {
// Records a printf-style message for zsc_last_error() and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
}  // namespace zsc
