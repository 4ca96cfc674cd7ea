// Per-thread error reporting shared by every C-ABI entry point.
#pragma once

#include <cstdarg>
#include <cstdio>

namespace indigo {

// Last error message of the calling host thread.
inline char* error_buffer()
{
  static thread_local char buf[512] = {0};
  return buf;
}

// Records a printf-style message and returns `code` so call sites can write
// `return set_error(INDIGO_ERR_..., "...")`.
inline int set_error(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(error_buffer(), 512, fmt, ap);
  va_end(ap);
  return code;
}

inline void clear_error() { error_buffer()[0] = '\0'; }

}  // namespace indigo
