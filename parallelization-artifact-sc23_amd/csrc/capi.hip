// extern "C" entry points: the drop-in entries (one per reference entry
// function) and the resident layer.  Argument checking, device selection,
// timing and copies live here; the kernels live in <alg>.hip.
#include <chrono>
#include <climits>

#include "bfs.hpp"
#include "runtime.hpp"

using namespace indigo;

namespace {

double wall_seconds()
{
  using clock = std::chrono::steady_clock;
  return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

// Device buffer released on scope exit (the drop-in entries own their
// temporaries, as the reference's entry functions do).
struct DeviceBuffer {
  void* p = nullptr;
  ~DeviceBuffer()
  {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes)
  {
    if (hipMalloc(&p, bytes > 0 ? bytes : 4) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return set_error(INDIGO_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes failed", bytes);
    }
    return INDIGO_OK;
  }
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(p);
  }
};

int ctx_prologue(indigo_ctx* ctx)
{
  clear_error();
  if (ctx == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "null context");
  IG_HIP_TRY(hipSetDevice(ctx->device));
  return INDIGO_OK;
}

int legacy_ctx(indigo_ctx** out)
{
  clear_error();
  int rc = INDIGO_OK;
  indigo_ctx* ctx = default_ctx(&rc);
  if (ctx == nullptr) return rc;
  IG_HIP_TRY(hipSetDevice(ctx->device));
  *out = ctx;
  return INDIGO_OK;
}

int check_src(int src, int nodes)
{
  if (src < 0 || src >= nodes)
    return set_error(INDIGO_ERR_INVALID_ARG, "source_node_number must be between 0 and %d (got %d)", nodes, src);
  return INDIGO_OK;
}

// ---- BFS -------------------------------------------------------------------
int bfs_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int src, int* dist,
               double* seconds)
{
  IG_TRY(ensure_workspace(ctx, bfs_workspace_bytes(nodes)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  IG_TRY(timing_begin(ctx));
  BfsStats st;
  IG_TRY(bfs_run(ctx, nodes, edges, nindex, nlist, src, dist, &st));
  IG_TRY(timing_end(ctx, st.launches, 0));
  IG_TRY(bfs_collect(ctx, &st));
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = st.levels;
  return INDIGO_OK;
}

}  // namespace

extern "C" {

int indigo_bfs(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int src,
               int* d_dist)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (d_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "bfs: dist is NULL");
  return bfs_device(ctx, nodes, edges, d_nindex, d_nlist, src, d_dist, nullptr);
}

int indigo_bfs_vertex(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* h_dist,
                      double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (h_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "bfs_vertex: dist is NULL");
  DeviceBuffer d;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(bfs_device(ctx, nodes, edges, d_nindex, d_nlist, src, d.as<int>(), seconds));
  IG_HIP_TRY(hipMemcpy(h_dist, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_bfs_edge(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* h_sp,
                    int* h_dist, double* seconds)
{
  if (h_sp == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "bfs_edge: sp is NULL");
  return indigo_bfs_vertex(src, nodes, edges, d_nindex, d_nlist, h_dist, seconds);
}

}  // extern "C"
