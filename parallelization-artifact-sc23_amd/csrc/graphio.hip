// Graph uploads and device-side edge sources (MI355X, gfx950).
//
// indigo_graph_upload moves a host CSR (malloc'd, or an .egr file mapped by
// indigo_egr_open) into new device arrays through two pinned staging
// buffers: while the DMA engine drains one, host threads fill the other, so
// the PCIe copy overlaps the page-in and memcpy of the next chunk.  A
// pageable hipMemcpy would stage through the runtime's own small buffers on
// one thread.
//
// indigo_build_sp writes the edge-source array sp[j] = v for every CSR entry
// j of vertex v on the device (the reference builds it on the host in every
// edge-form harness, e.g. include/bfs_edge_cuda.h:140-146, and copies E ints
// over PCIe).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "device_common.hpp"
#include "errors.hpp"
#include "graphio.hpp"
#include "runtime.hpp"

using namespace indigo;

namespace {

constexpr size_t kStageBytes = size_t(64) << 20;  // per staging buffer
constexpr size_t kCopyThreadBytes = size_t(8) << 20;

// Host memcpy split over threads (a single thread reaches about a third of
// what PCIe Gen5 x16 takes).
void parallel_copy(void* dst, const void* src, size_t bytes)
{
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<size_t>(hw, 8), std::max<size_t>(1, bytes / kCopyThreadBytes));
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> ts;
  const size_t per = (bytes + nt - 1) / nt;
  for (size_t t = 0; t < nt; t++) {
    const size_t b = t * per, e = std::min(bytes, b + per);
    if (b >= e) break;
    ts.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b); });
  }
  for (auto& t : ts) t.join();
}

int ensure_stage(indigo_ctx* ctx)
{
  if (ctx->stage_bytes >= kStageBytes) return INDIGO_OK;
  for (int b = 0; b < 2; b++) {
    if (ctx->stage[b] == nullptr) IG_HIP_TRY(hipHostMalloc(&ctx->stage[b], kStageBytes, hipHostMallocDefault));
    if (ctx->ev_stage[b] == nullptr) IG_HIP_TRY(hipEventCreateWithFlags(&ctx->ev_stage[b], hipEventDisableTiming));
  }
  ctx->stage_bytes = kStageBytes;
  return INDIGO_OK;
}

// Copies `bytes` from host `src` to device `dst` through the staging pair;
// `turn` alternates the buffers across calls of one upload.
int staged_copy(indigo_ctx* ctx, void* dst, const void* src, size_t bytes, int* turn)
{
  size_t off = 0;
  while (off < bytes) {
    const int b = *turn & 1;
    const size_t n = std::min(kStageBytes, bytes - off);
    // the DMA that last read this buffer must be done before it is refilled
    IG_HIP_TRY(hipEventSynchronize(ctx->ev_stage[b]));
    parallel_copy(ctx->stage[b], static_cast<const char*>(src) + off, n);
    IG_HIP_TRY(hipMemcpyAsync(static_cast<char*>(dst) + off, ctx->stage[b], n, hipMemcpyHostToDevice, ctx->stream));
    IG_HIP_TRY(hipEventRecord(ctx->ev_stage[b], ctx->stream));
    off += n;
    (*turn)++;
  }
  return INDIGO_OK;
}

// sp[j] = the vertex v with nindex[v] <= j < nindex[v + 1]: the largest v
// with nindex[v] <= j (vertices without entries before it share its start,
// those after it start beyond j).  One thread per entry, binary search over
// nindex (16.8 MB at R-MAT scale 22: the upper levels stay in L2).
__global__ __launch_bounds__(256) void sp_kernel(const int* __restrict__ nindex, int nodes, int edges,
                                                 int* __restrict__ sp)
{
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; j < edges; j += stride) {
    int lo = 0, hi = nodes - 1;
    while (lo < hi) {
      const int mid = lo + ((hi - lo + 1) >> 1);
      if (nindex[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    sp[j] = lo;
  }
}

// out[v] = in[perm[v]], 4-byte values.
__global__ __launch_bounds__(256) void unpermute_kernel(int n, const int* __restrict__ perm, const int* __restrict__ in,
                                                        int* __restrict__ out)
{
  const int stride = gridDim.x * blockDim.x;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) out[v] = in[perm[v]];
}

}  // namespace

int indigo::build_sp_async(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, int* d_sp)
{
  if (edges <= 0) return INDIGO_OK;
  const int block = 256;
  const long long want = (static_cast<long long>(edges) + block - 1) / block;
  const int grid = static_cast<int>(std::min<long long>(want, 32LL * ctx->compute_units));
  sp_kernel<<<grid, block, 0, ctx->stream>>>(d_nindex, nodes, edges, d_sp);
  IG_HIP_TRY(hipGetLastError());
  return INDIGO_OK;
}

extern "C" {

int indigo_graph_upload(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist,
                        const int* eweight, indigo_graph_ref* out, double* seconds)
{
  clear_error();
  if (ctx == nullptr || out == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "graph_upload: null argument");
  *out = indigo_graph_ref{};
  IG_TRY(check_graph_args(nodes, edges, nindex, nlist));
  IG_HIP_TRY(hipSetDevice(ctx->device));
  IG_TRY(ensure_stage(ctx));
  const size_t bi = sizeof(int) * (static_cast<size_t>(nodes) + 1), be = sizeof(int) * static_cast<size_t>(edges);
  void *d_i = nullptr, *d_l = nullptr, *d_w = nullptr;
  auto fail = [&](int rc) {
    if (d_i) (void)hipFree(d_i);
    if (d_l) (void)hipFree(d_l);
    if (d_w) (void)hipFree(d_w);
    return rc;
  };
  if (hipMalloc(&d_i, bi) != hipSuccess || hipMalloc(&d_l, std::max<size_t>(be, 4)) != hipSuccess ||
      (eweight != nullptr && hipMalloc(&d_w, std::max<size_t>(be, 4)) != hipSuccess))
    return fail(set_error(INDIGO_ERR_OUT_OF_MEMORY, "graph_upload: device allocation of %zu bytes failed",
                          bi + be * (eweight ? 2 : 1)));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const auto t0 = std::chrono::steady_clock::now();
  int turn = 0, rc = staged_copy(ctx, d_i, nindex, bi, &turn);
  if (rc == INDIGO_OK) rc = staged_copy(ctx, d_l, nlist, be, &turn);
  if (rc == INDIGO_OK && eweight != nullptr) rc = staged_copy(ctx, d_w, eweight, be, &turn);
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (rc != INDIGO_OK) return fail(rc);
  if (e != hipSuccess) return fail(set_error(INDIGO_ERR_DEVICE, "graph_upload: %s", hipGetErrorString(e)));
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = indigo_graph_ref{nodes, edges, static_cast<const int*>(d_i), static_cast<const int*>(d_l),
                          static_cast<const int*>(d_w)};
  return INDIGO_OK;
}

int indigo_graph_free(indigo_graph_ref* g)
{
  if (g == nullptr) return INDIGO_OK;
  if (g->d_nindex) (void)hipFree(const_cast<int*>(g->d_nindex));
  if (g->d_nlist) (void)hipFree(const_cast<int*>(g->d_nlist));
  if (g->d_eweight) (void)hipFree(const_cast<int*>(g->d_eweight));
  *g = indigo_graph_ref{};
  return INDIGO_OK;
}

int indigo_build_sp(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, int* d_sp)
{
  clear_error();
  if (ctx == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "build_sp: null context");
  IG_TRY(check_graph_args(nodes, edges, d_nindex, d_sp));
  IG_HIP_TRY(hipSetDevice(ctx->device));
  if (edges == 0) return INDIGO_OK;
  IG_TRY(timing_begin(ctx));
  IG_TRY(kernel_begin(ctx, "sp_kernel"));
  IG_TRY(build_sp_async(ctx, nodes, edges, d_nindex, d_sp));
  IG_TRY(kernel_end(ctx));
  return timing_end(ctx, 1, 1);
}

int indigo_unpermute(indigo_ctx* ctx, int n, const int* d_perm, const void* d_in, void* d_out)
{
  clear_error();
  if (ctx == nullptr || n < 0 || (n > 0 && (d_perm == nullptr || d_in == nullptr || d_out == nullptr)) ||
      (n > 0 && d_in == d_out))
    return set_error(INDIGO_ERR_INVALID_ARG, "unpermute: bad arguments");
  IG_HIP_TRY(hipSetDevice(ctx->device));
  if (n == 0) return INDIGO_OK;
  IG_TRY(timing_begin(ctx));
  IG_TRY(kernel_begin(ctx, "unpermute_kernel"));
  const int grid = static_cast<int>(std::min<long long>((static_cast<long long>(n) + 255) / 256, 8LL * ctx->compute_units));
  unpermute_kernel<<<grid, 256, 0, ctx->stream>>>(n, d_perm, static_cast<const int*>(d_in), static_cast<int*>(d_out));
  IG_HIP_TRY(hipGetLastError());
  IG_TRY(kernel_end(ctx));
  return timing_end(ctx, 1, 1);
}

}  // extern "C"
