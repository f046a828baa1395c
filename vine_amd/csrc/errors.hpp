// Thread-local error message behind vine_last_error() (include/vine.h).
#pragma once

namespace vine {

// Formats the message returned by vine_last_error() on this thread.
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

}  // namespace vine
