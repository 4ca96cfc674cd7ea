// Host side of the per-level traces: copies a kernel's LevelTrace into the
// context (microseconds since the kernel started) for indigo_ctx_last_trace.
#pragma once

#include "frontier.hpp"
#include "runtime.hpp"

namespace indigo {

inline void store_trace(indigo_ctx* ctx, const LevelTrace& tr)
{
  const int n = tr.n < kTraceMax ? tr.n : kTraceMax;
  ctx->trace_n = n;
  const double us_per_tick = 1e6 / static_cast<double>(ctx->wall_hz);
  for (int i = 0; i < n; i++) {
    ctx->trace_us[i] = static_cast<float>((tr.t[i] - tr.start) * us_per_tick);
    ctx->trace_nf[i] = tr.nf[i];
    ctx->trace_m[i] = static_cast<double>(tr.m[i]);
    ctx->trace_mode[i] = tr.mode[i];
  }
}

}  // namespace indigo
