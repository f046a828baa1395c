// Thread-local error message shared by every entry point of the library.
#include "errors.hpp"

#include <cstdarg>
#include <cstdio>
#include <string>

#include "vine.h"

namespace {
thread_local std::string g_err;
}

namespace vine {

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

}  // namespace vine

extern "C" const char* vine_last_error(void) { return g_err.c_str(); }
