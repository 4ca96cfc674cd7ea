// Host-side runtime shared by the kernels' launchers: the context object
// behind indigo_ctx, workspace carving, HIP error mapping and persistent
// (persistent) launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "errors.hpp"
#include "grid_plan.hpp"
#include "indigo_hip.h"

struct indigo_ctx {
  int device = 0;
  int compute_units = 0;
  int max_threads_per_cu = 0;
  long long wall_hz = 100000000;  // wall_clock64() frequency
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // events around the main kernel launch of the last call (indigo_ctx_last_kernel)
  hipEvent_t ev_k0 = nullptr, ev_k1 = nullptr;
  const char* kernel_name = nullptr;  // set by kernel_begin, a string literal
  float last_kernel_ms = 0.f;
  const char* last_kernel = "";
  void* ws = nullptr;      // cached device workspace
  size_t ws_bytes = 0;
  void* host_pinned = nullptr;  // small pinned staging buffer (readbacks)
  size_t host_pinned_bytes = 0;
  float last_ms = 0.f;
  int last_launches = 0;
  int last_iterations = 0;
  int hints[16] = {0};
  double budget_seconds = 60.0;  // bound on any device-side wait
  void* diag[4] = {nullptr, nullptr, nullptr, nullptr};  // device diagnostics of the last call
  // per-level trace of the last traversal call (see indigo_ctx_last_trace)
  int trace_n = 0;
  float trace_us[256] = {0};
  unsigned trace_nf[256] = {0};
  double trace_m[256] = {0};
  int trace_mode[256] = {0};
  // batch launched by batch_launch and not yet waited for
  struct {
    bool active = false;
    void* tasks = nullptr;
    void* gangs = nullptr;
    void* ctl = nullptr;
    int ntasks = 0;
    int grid_cap = 0;
  } pending_batch;
  // pinned double buffer of staged graph uploads (indigo_graph_upload)
  void* stage[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;
  hipEvent_t ev_stage[2] = {nullptr, nullptr};
  int last_grid = 0;    // workgroups of the last persistent launch
  int last_per_cu = 0;  // co-resident workgroups per CU it was sized for
  // grid cap per persistent kernel, learned from grids of this context that
  // did not assemble (check_abort); it ends with the context
  std::unordered_map<const void*, int> grid_caps;
  // kernel sequences captured into HIP graphs (INDIGO_HINT_GRAPH), keyed by
  // everything the sequence depends on; they refer to the workspace, so any
  // change of it destroys them
  struct Captured {
    unsigned long long key;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
  bool capturing = false;  // a sequence is being captured: no event records
};

namespace indigo {

#define IG_HIP_TRY(expr)                                                                       \
  do {                                                                                         \
    hipError_t ig_e_ = (expr);                                                                 \
    if (ig_e_ != hipSuccess)                                                                   \
      return ::indigo::set_error(INDIGO_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr,           \
                                 hipGetErrorString(ig_e_), __FILE__, __LINE__);                \
  } while (0)

#define IG_TRY(expr)                    \
  do {                                  \
    int ig_rc_ = (expr);                \
    if (ig_rc_ != INDIGO_OK) return ig_rc_; \
  } while (0)

// Grows the context workspace to at least `bytes` (contents not preserved).
int ensure_workspace(indigo_ctx* ctx, size_t bytes);
int ensure_host_pinned(indigo_ctx* ctx, size_t bytes);

// Sequentially carves 256-byte-aligned sub-buffers out of a workspace.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* p) : base(static_cast<char*>(p)) {}
  template <typename T>
  T* take(size_t count)
  {
    off = (off + 255) & ~static_cast<size_t>(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += sizeof(T) * (count > 0 ? count : 1);
    return p;
  }
  static size_t round(size_t bytes) { return (bytes + 255) & ~static_cast<size_t>(255); }
};

// Grid of an ordinary (grid-stride or work-queue) launch within the
// context's share of the GPU: INDIGO_HINT_MAX_WORKGROUPS counts workgroups
// of 512 threads, so a grid of `block`-thread workgroups is held to as many
// threads.  Without it, one context's full-GPU grid (TC's counting kernels
// hold every CU until their queue drains) keeps the persistent grids of the
// contexts beside it from assembling.
inline int share_grid(const indigo_ctx* ctx, int grid, int block)
{
  const int cap = ctx->hints[INDIGO_HINT_MAX_WORKGROUPS];
  if (cap <= 0 || block <= 0) return grid;
  const long long t = static_cast<long long>(cap) * 512 / block;
  return static_cast<int>(std::max<long long>(1, std::min<long long>(grid, t)));
}

// Ticks of wall_clock64() that a persistent kernel may spend in bounded waits.
inline long long budget_ticks(const indigo_ctx* ctx)
{
  const double sec = ctx->hints[INDIGO_HINT_BUDGET_MS] > 0 ? ctx->hints[INDIGO_HINT_BUDGET_MS] * 1e-3 : ctx->budget_seconds;
  return static_cast<long long>(sec * static_cast<double>(ctx->wall_hz));
}

// Launches a persistent `kernel` with as many workgroups of `block` threads
// as can be co-resident: occupancy x compute units, capped at max_blocks, by
// the context's hints and by the context's cap for this kernel learned from
// earlier launches (plan_grid in grid_plan.hpp).  It is a
// plain launch: the kernels' grid barrier is software (device atomics), and
// co-residency is checked by the kernel itself (residency_barrier), which
// aborts a grid that does not assemble so the host relaunches it smaller.
// (A cooperative launch would add nothing, and on ROCm it fails while
// another process holds a context on the same GPU.)
// `name` (a string literal, or null) marks the launch as the call's main
// kernel, timed by kernel_begin / kernel_end.  full_cu: the kernel is held
// to 80 scalar registers (IG_SGPR_CAP) and may fill every wave slot of a CU.
int coop_launch(indigo_ctx* ctx, const void* kernel, int block, size_t shmem, void** args, int max_blocks,
                int* grid_out, const char* name = nullptr, bool full_cu = false);
// Workgroups of `kernel` that coop_launch places per CU before any learned
// cap or hint (occupancy, less one when that would fill the CU unless
// full_cu).
int resident_per_cu(indigo_ctx* ctx, const void* kernel, int block, size_t shmem, int* per_cu, bool full_cu = false);
// Internal status: the last persistent launch did not assemble (see
// residency_barrier); the context's cap of `kernel` has been lowered below
// the grid actually launched (lowered_cap in grid_plan.hpp), launch again.
constexpr int kNotResident = -100;
// Checks a finished persistent kernel's barrier abort code: 0 ok,
// kNotResident (cap lowered), or INDIGO_ERR_TIMEOUT.
int check_abort(indigo_ctx* ctx, const void* kernel, unsigned abort_code, const char* what);
// Ticks a persistent kernel allows for its grid to assemble.
inline long long residency_ticks(const indigo_ctx* ctx)
{
  const int ms = ctx->hints[INDIGO_HINT_RESIDENCY_MS];
  return ms > 0 ? ctx->wall_hz / 1000 * ms : ctx->wall_hz / 10;
}

// Enqueues the kernel sequence `enqueue(ctx)` on ctx->stream.  With
// INDIGO_HINT_GRAPH set, the sequence is captured into a HIP graph the first
// time a call with this `key` comes (the key must cover every argument,
// pointer and hint the sequence depends on) and the graph is replayed
// afterwards: one launch instead of one per kernel.  The sequence must not
// synchronise or copy to the host.
int run_captured(indigo_ctx* ctx, unsigned long long key, int (*enqueue)(indigo_ctx*, void*), void* arg);
// Destroys the context's captured graphs.
void drop_captured(indigo_ctx* ctx);
// Mixes one more word into a graph key.
inline unsigned long long key_mix(unsigned long long h, unsigned long long v)
{
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xff51afd7ed558ccdull;
}

// Zeroes `bytes` of device memory with a kernel (no memset node, so the
// sequence stays safe to capture into a HIP graph).
int zero_async(indigo_ctx* ctx, void* p, size_t bytes);

// Argument checks shared by the entry points.
int check_graph_args(int nodes, int edges, const void* nindex, const void* nlist);

// Starts / stops the HIP-event timing of a resident call.
int timing_begin(indigo_ctx* ctx);
// timing_end = timing_record_stop + timing_finish: the first records the stop
// event without waiting, the second waits for it and stores the times.
int timing_record_stop(indigo_ctx* ctx);
int timing_finish(indigo_ctx* ctx, int launches, int iterations);
// Events around the main kernel of a call (the last marked launch wins);
// `name` must be a string literal.
int kernel_begin(indigo_ctx* ctx, const char* name);
int kernel_end(indigo_ctx* ctx);
int timing_end(indigo_ctx* ctx, int launches, int iterations);

// Host wall clock in seconds (the drop-in entries' timer).
double wall_seconds();

// Process-wide default context for the drop-in entries (created lazily on
// device 0, or on the current device); indigo_shutdown releases it.
indigo_ctx* default_ctx(int* rc);

}  // namespace indigo
