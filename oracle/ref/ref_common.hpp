// Shared helpers of the wrapper translation units that compile the
// reference's own C++ sources (compiled where they lie, never copied) into
// oracle/_ref/libref_<alg>.so.  Test infrastructure only.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <unistd.h>

// Runs fn() with stdout redirected to a temporary file and returns the
// captured text (malloc'd, NUL-terminated); the reference prints its
// iteration counts instead of returning them.
template <typename F>
static char* ref_capture_stdout(F fn)
{
  fflush(stdout);
  FILE* tmp = tmpfile();
  const int saved = dup(1);
  dup2(fileno(tmp), 1);
  fn();
  fflush(stdout);
  dup2(saved, 1);
  close(saved);
  const long n = ftell(tmp);
  char* buf = static_cast<char*>(malloc(static_cast<size_t>(n) + 1));
  rewind(tmp);
  const size_t got = fread(buf, 1, static_cast<size_t>(n), tmp);
  buf[got] = '\0';
  fclose(tmp);
  return buf;
}
