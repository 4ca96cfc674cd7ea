// extern "C" entry points: the drop-in entries (one per reference entry
// function) and the resident layer.  Argument checking, device selection,
// timing and copies live here; the kernels live in <alg>.hip.
#include <chrono>
#include <climits>

#include "batch.hpp"
#include "bfs.hpp"
#include "server.hpp"
#include "cc.hpp"
#include "device_common.hpp"
#include "edge.hpp"
#include "mis.hpp"
#include "pr.hpp"
#include "runtime.hpp"
#include "sssp.hpp"
#include "tc.hpp"

using namespace indigo;

namespace {

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

// INDIGO_HINT_VARIANT selects the edge-centric form for the resident calls.
bool edge_variant(const indigo_ctx* ctx) { return ctx->hints[INDIGO_HINT_VARIANT] == INDIGO_VARIANT_EDGE; }

// Copies the harness's host edge-source array to the device (the reference's
// edge forms copy sp before their timer starts, e.g.
// codes/cuda/bfs-cuda/BFS_CUDA_E_Topo_*.cu).
int upload_sp(DeviceBuffer& d, const int* h_sp, int edges)
{
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(edges > 0 ? edges : 1)));
  if (edges > 0) IG_HIP_TRY(hipMemcpy(d.p, h_sp, sizeof(int) * static_cast<size_t>(edges), hipMemcpyHostToDevice));
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
  BfsStats st;
  for (;;) {
    IG_TRY(timing_begin(ctx));
    IG_TRY(bfs_run(ctx, nodes, edges, nindex, nlist, src, dist, &st));
    IG_TRY(timing_end(ctx, st.launches, 0));
    const int rc = bfs_collect(ctx, &st);
    if (rc != kNotResident) IG_TRY(rc);
    if (rc == INDIGO_OK) break;
  }
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = st.levels;
  return INDIGO_OK;
}


// ---- SSSP / CC / MIS / PR / TC on device buffers ------------------------------
int sssp_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, const int* eweight,
                int src, int* dist, double* seconds)
{
  IG_TRY(ensure_workspace(ctx, sssp_workspace_bytes(nodes)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  int iters = 0, epochs = 0;
  for (;;) {
    IG_TRY(timing_begin(ctx));
    IG_TRY(sssp_run(ctx, nodes, edges, nindex, nlist, eweight, src, dist));
    IG_TRY(timing_end(ctx, 2, 0));
    const int rc = sssp_collect(ctx, &iters, &epochs);
    if (rc != kNotResident) IG_TRY(rc);
    if (rc == INDIGO_OK) break;
  }
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = iters;
  return INDIGO_OK;
}

// Arguments of a captured kernel sequence (run_captured).
struct SeqArgs {
  int nodes, edges;
  const int* nindex;
  const int* nlist;
  int* out;
  int launches;
};

unsigned long long seq_key(const indigo_ctx* ctx, int alg, const SeqArgs& a)
{
  unsigned long long k = key_mix(0, static_cast<unsigned>(alg));
  k = key_mix(k, static_cast<unsigned>(a.nodes));
  k = key_mix(k, static_cast<unsigned>(a.edges));
  k = key_mix(k, reinterpret_cast<uintptr_t>(a.nindex));
  k = key_mix(k, reinterpret_cast<uintptr_t>(a.nlist));
  for (int h : ctx->hints) k = key_mix(k, static_cast<unsigned>(h));  // hints shape the launches
  return key_mix(k, reinterpret_cast<uintptr_t>(a.out));
}

int tc_sequence(indigo_ctx* ctx, void* p)
{
  SeqArgs* a = static_cast<SeqArgs*>(p);
  return tc_run(ctx, a->nodes, a->edges, a->nindex, a->nlist, &a->launches);
}

int cc_sequence(indigo_ctx* ctx, void* p)
{
  SeqArgs* a = static_cast<SeqArgs*>(p);
  return cc_run(ctx, a->nodes, a->edges, a->nindex, a->nlist, a->out, &a->launches);
}

int cc_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* label,
              double* seconds)
{
  IG_TRY(ensure_workspace(ctx, cc_workspace_bytes(nodes)));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  IG_TRY(timing_begin(ctx));
  SeqArgs a{nodes, edges, nindex, nlist, label, 0};
  IG_TRY(run_captured(ctx, seq_key(ctx, 3, a), cc_sequence, &a));
  IG_TRY(timing_end(ctx, a.launches, 1));
  if (seconds) *seconds = wall_seconds() - t0;
  return INDIGO_OK;
}

int mis_device(indigo_ctx* ctx, int nodes, const int* nindex, const int* nlist, int signed_priority, int* status,
               double* seconds)
{
  IG_TRY(ensure_workspace(ctx, mis_workspace_bytes(nodes)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  int rounds = 0;
  for (;;) {
    IG_TRY(timing_begin(ctx));
    IG_TRY(mis_run(ctx, nodes, nindex, nlist, signed_priority, status));
    IG_TRY(timing_end(ctx, 2, 0));
    const int rc = mis_collect(ctx, &rounds);
    if (rc != kNotResident) IG_TRY(rc);
    if (rc == INDIGO_OK) break;
  }
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = rounds;
  return INDIGO_OK;
}

int pr_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, const int* degree, float* scores,
              int init_uniform, int* iterations, double* seconds)
{
  IG_TRY(ensure_workspace(ctx, pr_workspace_bytes(nodes, edges)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  int iters = 0;
  for (;;) {
    IG_TRY(timing_begin(ctx));
    IG_TRY(pr_run(ctx, nodes, edges, nindex, nlist, degree, scores, init_uniform));
    IG_TRY(timing_end(ctx, 2, 0));
    const int rc = pr_collect(ctx, &iters, nullptr);
    if (rc != kNotResident) IG_TRY(rc);
    if (rc == INDIGO_OK) break;
  }
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = iters;
  if (iterations) *iterations = iters;
  return INDIGO_OK;
}

int tc_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, long long* count,
              double* seconds)
{
  IG_TRY(ensure_workspace(ctx, tc_workspace_bytes(nodes, edges)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  IG_TRY(timing_begin(ctx));
  SeqArgs a{nodes, edges, nindex, nlist, nullptr, 10};
  IG_TRY(run_captured(ctx, seq_key(ctx, 6, a), tc_sequence, &a));
  IG_TRY(timing_end(ctx, a.launches, 1));
  IG_TRY(tc_collect(ctx, count));
  if (seconds) *seconds = wall_seconds() - t0;
  return INDIGO_OK;
}

// Reads one int from device memory (used to derive sizes the reference's
// entry signatures do not pass).
int read_device_int(const int* d, int* out)
{
  IG_HIP_TRY(hipMemcpy(out, d, sizeof(int), hipMemcpyDeviceToHost));
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
  if (edge_variant(ctx))
    return edge_device(ctx, kEdgeBfs, nodes, edges, d_nindex, d_nlist, nullptr, nullptr, src, 0, d_dist, nullptr);
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
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (h_sp == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "bfs_edge: sp is NULL");
  if (h_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "bfs_edge: dist is NULL");
  DeviceBuffer d, dsp;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(upload_sp(dsp, h_sp, edges));
  IG_TRY(edge_device(ctx, kEdgeBfs, nodes, edges, d_nindex, d_nlist, nullptr, dsp.as<int>(), src, 0, d.as<int>(),
                     seconds));
  IG_HIP_TRY(hipMemcpy(h_dist, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_sssp(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                const int* d_eweight, int src, int* d_dist)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (d_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "sssp: dist is NULL");
  if (d_eweight == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "sssp: eweight is NULL");
  if (edge_variant(ctx))
    return edge_device(ctx, kEdgeSssp, nodes, edges, d_nindex, d_nlist, d_eweight, nullptr, src, 0, d_dist, nullptr);
  return sssp_device(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, src, d_dist, nullptr);
}

int indigo_sssp_vertex(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                       const int* d_eweight, int* h_dist, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (h_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "sssp_vertex: dist is NULL");
  if (d_eweight == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "sssp_vertex: eweight is NULL");
  DeviceBuffer d;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(sssp_device(ctx, nodes, edges, d_nindex, d_nlist, d_eweight, src, d.as<int>(), seconds));
  IG_HIP_TRY(hipMemcpy(h_dist, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_sssp_edge(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                     const int* d_eweight, const int* h_sp, int* h_dist, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  IG_TRY(check_src(src, nodes));
  if (h_sp == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "sssp_edge: sp is NULL");
  if (h_dist == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "sssp_edge: dist is NULL");
  if (d_eweight == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "sssp_edge: eweight is NULL");
  DeviceBuffer d, dsp;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(upload_sp(dsp, h_sp, edges));
  IG_TRY(edge_device(ctx, kEdgeSssp, nodes, edges, d_nindex, d_nlist, d_eweight, dsp.as<int>(), src, 0, d.as<int>(),
                     seconds));
  IG_HIP_TRY(hipMemcpy(h_dist, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_cc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (d_label == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "cc: label is NULL");
  if (edge_variant(ctx))
    return edge_device(ctx, kEdgeCc, nodes, edges, d_nindex, d_nlist, nullptr, nullptr, 0, 0, d_label, nullptr);
  return cc_device(ctx, nodes, edges, d_nindex, d_nlist, d_label, nullptr);
}

static_assert(INDIGO_CC_NOALLOC_MAX_NODES == kCcScratchMaxNodes, "scratch limit of the C header");

int indigo_cc_noalloc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (d_label == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "cc_noalloc: label is NULL");
  return cc_noalloc_device(ctx, nodes, edges, d_nindex, d_nlist, d_label);
}

int indigo_cc_vertex(int nodes, int edges, const int* d_nindex, const int* d_nlist, int* h_label, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (h_label == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "cc_vertex: label is NULL");
  DeviceBuffer d;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(cc_device(ctx, nodes, edges, d_nindex, d_nlist, d.as<int>(), seconds));
  IG_HIP_TRY(hipMemcpy(h_label, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_cc_edge(int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* h_sp, int* h_label,
                   double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (h_sp == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "cc_edge: sp is NULL");
  if (h_label == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "cc_edge: label is NULL");
  DeviceBuffer d, dsp;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(upload_sp(dsp, h_sp, edges));
  IG_TRY(edge_device(ctx, kEdgeCc, nodes, edges, d_nindex, d_nlist, nullptr, dsp.as<int>(), 0, 0, d.as<int>(), seconds));
  IG_HIP_TRY(hipMemcpy(h_label, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_mis(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int signed_priority,
               int* d_status)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (d_status == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "mis: status is NULL");
  if (edge_variant(ctx))
    return edge_device(ctx, kEdgeMis, nodes, edges, d_nindex, d_nlist, nullptr, nullptr, 0, signed_priority, d_status,
                       nullptr);
  return mis_device(ctx, nodes, d_nindex, d_nlist, signed_priority, d_status, nullptr);
}

int indigo_mis_vertex(int nodes, int edges, const int* d_nindex, const int* d_nlist, int signed_priority,
                      unsigned* h_priority, int* h_status, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (h_status == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "mis_vertex: status is NULL");
  DeviceBuffer d;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(mis_device(ctx, nodes, d_nindex, d_nlist, signed_priority, d.as<int>(), seconds));
  IG_HIP_TRY(hipMemcpy(h_status, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  if (h_priority != nullptr)
    for (int v = 0; v < nodes; v++) h_priority[v] = mis_hash(static_cast<unsigned>(v) + 712313887u);
  return INDIGO_OK;
}

int indigo_mis_edge(int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                    int signed_priority, unsigned* h_priority, int* h_status, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (d_sp == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "mis_edge: sp is NULL");
  if (h_status == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "mis_edge: status is NULL");
  DeviceBuffer d;
  IG_TRY(d.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(edge_device(ctx, kEdgeMis, nodes, edges, d_nindex, d_nlist, nullptr, d_sp, 0, signed_priority, d.as<int>(),
                     seconds));
  IG_HIP_TRY(hipMemcpy(h_status, d.p, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  if (h_priority != nullptr)
    for (int v = 0; v < nodes; v++) h_priority[v] = mis_hash(static_cast<unsigned>(v) + 712313887u);
  return INDIGO_OK;
}

int indigo_pagerank(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, float* d_scores,
                    int* iterations)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (d_scores == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "pagerank: scores is NULL");
  return pr_device(ctx, nodes, edges, d_nindex, d_nlist, nullptr, d_scores, 1, iterations, nullptr);
}

int indigo_pr(int nodes, int edges, const int* h_nindex, const int* h_nlist, const int* h_degree, float* h_scores,
              int* iterations, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  IG_TRY(check_graph_args(nodes, edges, h_nindex, h_nlist));
  if (h_scores == nullptr || h_degree == nullptr)
    return set_error(INDIGO_ERR_INVALID_ARG, "pr: scores or degree is NULL");
  // PR_GPU copies the host graph itself (include/pr_cuda.h:48, variant files)
  DeviceBuffer dn, dl, dd, ds;
  IG_TRY(dn.alloc(sizeof(int) * (static_cast<size_t>(nodes) + 1)));
  IG_TRY(dl.alloc(sizeof(int) * static_cast<size_t>(edges)));
  IG_TRY(dd.alloc(sizeof(int) * static_cast<size_t>(nodes)));
  IG_TRY(ds.alloc(sizeof(float) * static_cast<size_t>(nodes)));
  IG_HIP_TRY(hipMemcpy(dn.p, h_nindex, sizeof(int) * (static_cast<size_t>(nodes) + 1), hipMemcpyHostToDevice));
  if (edges > 0) IG_HIP_TRY(hipMemcpy(dl.p, h_nlist, sizeof(int) * static_cast<size_t>(edges), hipMemcpyHostToDevice));
  IG_HIP_TRY(hipMemcpy(dd.p, h_degree, sizeof(int) * static_cast<size_t>(nodes), hipMemcpyHostToDevice));
  IG_HIP_TRY(hipMemcpy(ds.p, h_scores, sizeof(float) * static_cast<size_t>(nodes), hipMemcpyHostToDevice));
  IG_TRY(pr_device(ctx, nodes, edges, dn.as<int>(), dl.as<int>(), dd.as<int>(), ds.as<float>(), 0, iterations, seconds));
  IG_HIP_TRY(hipMemcpy(h_scores, ds.p, sizeof(float) * static_cast<size_t>(nodes), hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_tc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, long long* count)
{
  IG_TRY(ctx_prologue(ctx));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  if (count == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "tc: count is NULL");
  return tc_device(ctx, nodes, edges, d_nindex, d_nlist, count, nullptr);
}

int indigo_tc_vertex(long long* count, int nodes, const int* d_nindex, const int* d_nlist, double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  if (count == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "tc_vertex: count is NULL");
  if (nodes < 1 || d_nindex == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "tc_vertex: bad graph");
  int edges = 0;
  IG_TRY(read_device_int(d_nindex + nodes, &edges));
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_nlist));
  return tc_device(ctx, nodes, edges, d_nindex, d_nlist, count, seconds);
}

int indigo_tc_edge(long long* count, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                   double* seconds)
{
  indigo_ctx* ctx = nullptr;
  IG_TRY(legacy_ctx(&ctx));
  if (count == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "tc_edge: count is NULL");
  if (edges < 0 || d_nindex == nullptr || (edges > 0 && (d_nlist == nullptr || d_sp == nullptr)))
    return set_error(INDIGO_ERR_INVALID_ARG, "tc_edge: bad graph");
  if (edges == 0) {
    *count = 0;
    if (seconds) *seconds = 0.0;
    return INDIGO_OK;
  }
  // sp is sorted (tc_edge_cuda.h:159-165): the last entry's source + 1 bounds
  // every vertex with a list; trailing isolated vertices hold no triangles
  int last = 0;
  IG_TRY(read_device_int(d_sp + edges - 1, &last));
  const int nodes = last + 1;
  return tc_device(ctx, nodes, edges, d_nindex, d_nlist, count, seconds);
}

int indigo_batch(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
                 int nqueries, int max_workgroups)
{
  IG_TRY(ctx_prologue(ctx));
  return batch_run(ctx, graphs, ngraphs, queries, nqueries, max_workgroups);
}

int indigo_batch_gang(int alg, int nodes, int edges) { return batch_gang(alg, nodes, edges); }

int indigo_serve(const indigo_graph_ref* graphs, int ngraphs, indigo_request* reqs, int nreqs,
                 const indigo_serve_config* cfg, indigo_serve_stats* stats)
{
  clear_error();
  if (cfg == nullptr || cfg->slots < 1 || cfg->slots > 16 || cfg->max_batch < 1 || cfg->mem_budget == 0)
    return set_error(INDIGO_ERR_INVALID_ARG, "serve: bad configuration");
  if (cfg->batch_work_us < 0) return set_error(INDIGO_ERR_INVALID_ARG, "serve: batch_work_us must be >= 0");
  if (cfg->cu_partition != 0)
    return set_error(INDIGO_ERR_INVALID_ARG, "serve: the CU partition between slots is switched off");
  if (nreqs < 0 || ngraphs < 1 || graphs == nullptr || (nreqs > 0 && reqs == nullptr))
    return set_error(INDIGO_ERR_INVALID_ARG, "serve: bad graph / request table");
  for (int g = 0; g < ngraphs; g++) IG_TRY(check_graph_args(graphs[g].nodes, graphs[g].edges, graphs[g].d_nindex,
                                                           graphs[g].d_nlist));
  for (int i = 0; i < nreqs; i++) {
    const indigo_request& r = reqs[i];
    if (r.graph < 0 || r.graph >= ngraphs) return set_error(INDIGO_ERR_INVALID_ARG, "serve: request %d graph", i);
    const indigo_graph_ref& G = graphs[r.graph];
    switch (r.alg) {
      case INDIGO_QUERY_SSSP:
        if (G.d_eweight == nullptr && G.edges > 0)
          return set_error(INDIGO_ERR_INVALID_ARG, "serve: request %d is SSSP on a graph without weights", i);
        [[fallthrough]];
      case INDIGO_QUERY_BFS:
        if (r.src < 0 || r.src >= G.nodes) return set_error(INDIGO_ERR_INVALID_ARG, "serve: request %d source", i);
        break;
      case INDIGO_QUERY_CC:
      case INDIGO_QUERY_MIS:
      case INDIGO_QUERY_PR:
        break;
      case INDIGO_QUERY_TC:
        if (G.edges == 0) return set_error(INDIGO_ERR_INVALID_ARG, "serve: request %d is TC on an empty graph", i);
        break;
      default:
        return set_error(INDIGO_ERR_INVALID_ARG, "serve: request %d has unknown algorithm %d", i, r.alg);
    }
  }
  if (nreqs == 0) return INDIGO_OK;
  IG_HIP_TRY(hipSetDevice(cfg->device));
  return serve_run(graphs, ngraphs, reqs, nreqs, cfg, stats);
}

}  // extern "C"
