// Context management, workspace, persistent launches and device plumbing.
#include "runtime.hpp"

#include <cstdio>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "device_common.hpp"

namespace indigo {

namespace {

__global__ void zero_words_kernel(unsigned int* p, size_t words)
{
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < words;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

}  // namespace

void drop_captured(indigo_ctx* ctx)
{
  for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
}

int run_captured(indigo_ctx* ctx, unsigned long long key, int (*enqueue)(indigo_ctx*, void*), void* arg)
{
  if (ctx->hints[INDIGO_HINT_GRAPH] == 0) return enqueue(ctx, arg);
  key = key_mix(key, reinterpret_cast<uintptr_t>(ctx->ws));
  key = key_mix(key, ctx->ws_bytes);
  for (int h = 0; h < 16; h++) key = key_mix(key, static_cast<unsigned>(ctx->hints[h]));
  for (const auto& g : ctx->graphs)
    if (g.key == key) {
      IG_HIP_TRY(hipGraphLaunch(g.exec, ctx->stream));
      return INDIGO_OK;
    }
  IG_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  const int rc = enqueue(ctx, arg);
  ctx->capturing = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc != INDIGO_OK) {
    if (e == hipSuccess && graph != nullptr) (void)hipGraphDestroy(graph);
    return rc;
  }
  IG_HIP_TRY(e);
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  IG_HIP_TRY(ie);
  if (ctx->graphs.size() >= 16) {  // keep the 16 most recent sequences
    (void)hipGraphExecDestroy(ctx->graphs.front().exec);
    ctx->graphs.erase(ctx->graphs.begin());
  }
  ctx->graphs.push_back(indigo_ctx::Captured{key, exec});
  IG_HIP_TRY(hipGraphLaunch(exec, ctx->stream));
  return INDIGO_OK;
}

int ensure_workspace(indigo_ctx* ctx, size_t bytes)
{
  if (bytes <= ctx->ws_bytes && ctx->ws != nullptr) return INDIGO_OK;
  if (ctx->ws != nullptr) {
    IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    drop_captured(ctx);
    IG_HIP_TRY(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  const size_t want = Carver::round(std::max<size_t>(bytes, 1 << 20));
  if (hipMalloc(&ctx->ws, want) != hipSuccess) {
    (void)hipGetLastError();
    ctx->ws = nullptr;
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "workspace allocation of %zu bytes failed", want);
  }
  ctx->ws_bytes = want;
  return INDIGO_OK;
}

int ensure_host_pinned(indigo_ctx* ctx, size_t bytes)
{
  if (bytes <= ctx->host_pinned_bytes && ctx->host_pinned != nullptr) return INDIGO_OK;
  if (ctx->host_pinned != nullptr) {
    IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    IG_HIP_TRY(hipHostFree(ctx->host_pinned));
    ctx->host_pinned = nullptr;
    ctx->host_pinned_bytes = 0;
  }
  const size_t want = std::max<size_t>(bytes, 4096);
  if (hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    ctx->host_pinned = nullptr;
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "pinned host allocation of %zu bytes failed", want);
  }
  ctx->host_pinned_bytes = want;
  return INDIGO_OK;
}

int resident_per_cu(indigo_ctx* ctx, const void* kernel, int block, size_t shmem, int* per_cu, bool full_cu)
{
  int n = 0;
  IG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, shmem));
  if (n < 1) return set_error(INDIGO_ERR_DEVICE, "kernel cannot be resident with block %d", block);
  *per_cu = plan_per_cu(n, block, ctx->max_threads_per_cu, full_cu);
  return INDIGO_OK;
}

int coop_launch(indigo_ctx* ctx, const void* kernel, int block, size_t shmem, void** args, int max_blocks,
                int* grid_out, const char* name, bool full_cu)
{
  int per_cu = 0;
  IG_TRY(resident_per_cu(ctx, kernel, block, shmem, &per_cu, full_cu));
  GridLimits lim;
  lim.compute_units = ctx->compute_units;
  lim.hint_blocks_per_cu = ctx->hints[INDIGO_HINT_BLOCKS_PER_CU];
  lim.max_blocks = max_blocks;
  lim.hint_max_workgroups = ctx->hints[INDIGO_HINT_MAX_WORKGROUPS];
  const auto it = ctx->grid_caps.find(kernel);
  lim.learned_cap = it != ctx->grid_caps.end() ? it->second : 0;
  const int grid = plan_grid(per_cu, lim);
  if (name != nullptr) IG_TRY(kernel_begin(ctx, name));
  IG_HIP_TRY(hipLaunchKernel(kernel, dim3(grid), dim3(block), args, static_cast<unsigned>(shmem), ctx->stream));
  if (name != nullptr) IG_TRY(kernel_end(ctx));
  ctx->last_grid = grid;
  ctx->last_per_cu = lim.hint_blocks_per_cu > 0 ? std::min(per_cu, lim.hint_blocks_per_cu) : per_cu;
  if (grid_out) *grid_out = grid;
  return INDIGO_OK;
}

int check_abort(indigo_ctx* ctx, const void* kernel, unsigned abort_code, const char* what)
{
  if (abort_code == 0u) return INDIGO_OK;
  if (abort_code == kAbortNotResident) {
    const int cap = lowered_cap(ctx->last_grid, ctx->compute_units);
    if (cap < 1)
      return set_error(INDIGO_ERR_DEVICE, "%s: a grid of %d workgroup never assembled", what, ctx->last_grid);
    ctx->grid_caps[kernel] = cap;
    return kNotResident;
  }
  return set_error(INDIGO_ERR_TIMEOUT, "%s: persistent kernel stopped at a bounded wait (abort %u)", what,
                   abort_code);
}

double wall_seconds()
{
  using clock = std::chrono::steady_clock;
  return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

int zero_async(indigo_ctx* ctx, void* p, size_t bytes)
{
  const size_t words = (bytes + 3) / 4;
  if (words == 0) return INDIGO_OK;
  const int block = 256;
  const int grid = static_cast<int>(std::min<size_t>((words + block - 1) / block, 4096));
  zero_words_kernel<<<grid, block, 0, ctx->stream>>>(static_cast<unsigned int*>(p), words);
  IG_HIP_TRY(hipGetLastError());
  return INDIGO_OK;
}

int check_graph_args(int nodes, int edges, const void* nindex, const void* nlist)
{
  if (nodes < 1) return set_error(INDIGO_ERR_INVALID_ARG, "nodes must be >= 1 (got %d)", nodes);
  if (edges < 0) return set_error(INDIGO_ERR_INVALID_ARG, "edges must be >= 0 (got %d)", edges);
  if (nindex == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "nindex is NULL");
  if (nlist == nullptr && edges > 0) return set_error(INDIGO_ERR_INVALID_ARG, "nlist is NULL");
  return INDIGO_OK;
}

int timing_begin(indigo_ctx* ctx)
{
  ctx->kernel_name = nullptr;
  IG_HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
  return INDIGO_OK;
}

int kernel_begin(indigo_ctx* ctx, const char* name)
{
  if (ctx->capturing) return INDIGO_OK;  // a replayed graph is timed as a whole
  IG_HIP_TRY(hipEventRecord(ctx->ev_k0, ctx->stream));
  ctx->kernel_name = name;
  return INDIGO_OK;
}

int kernel_end(indigo_ctx* ctx)
{
  if (ctx->capturing) return INDIGO_OK;
  IG_HIP_TRY(hipEventRecord(ctx->ev_k1, ctx->stream));
  return INDIGO_OK;
}

int timing_record_stop(indigo_ctx* ctx)
{
  IG_HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
  return INDIGO_OK;
}

int timing_finish(indigo_ctx* ctx, int launches, int iterations)
{
  IG_HIP_TRY(hipEventSynchronize(ctx->ev_stop));
  float ms = 0.f;
  IG_HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
  ctx->last_ms = ms;
  if (ctx->kernel_name != nullptr) {
    float kms = 0.f;
    IG_HIP_TRY(hipEventElapsedTime(&kms, ctx->ev_k0, ctx->ev_k1));
    ctx->last_kernel_ms = kms;
    ctx->last_kernel = ctx->kernel_name;
  } else {
    ctx->last_kernel_ms = ms;
    ctx->last_kernel = "";
  }
  ctx->last_launches = launches;
  ctx->last_iterations = iterations;
  return INDIGO_OK;
}

int timing_end(indigo_ctx* ctx, int launches, int iterations)
{
  IG_TRY(timing_record_stop(ctx));
  return timing_finish(ctx, launches, iterations);
}

static int ctx_init(int device, indigo_ctx* c)
{
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return set_error(INDIGO_ERR_NO_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= count) return set_error(INDIGO_ERR_INVALID_ARG, "device %d out of range", device);
  IG_HIP_TRY(hipSetDevice(device));
  c->device = device;
  hipDeviceProp_t prop;
  IG_HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->compute_units = prop.multiProcessorCount;
  c->max_threads_per_cu = prop.maxThreadsPerMultiProcessor;
  int wall_khz = 0;
  if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wall_khz > 0)
    c->wall_hz = static_cast<long long>(wall_khz) * 1000;
  IG_HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  IG_HIP_TRY(hipEventCreate(&c->ev_start));
  IG_HIP_TRY(hipEventCreate(&c->ev_stop));
  IG_HIP_TRY(hipEventCreate(&c->ev_k0));
  IG_HIP_TRY(hipEventCreate(&c->ev_k1));
  return INDIGO_OK;
}

namespace {
std::mutex g_default_mu;
indigo_ctx* g_default = nullptr;  // the drop-in entries' context
}  // namespace

indigo_ctx* default_ctx(int* rc)
{
  std::lock_guard<std::mutex> lock(g_default_mu);
  if (g_default == nullptr) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
      (void)hipGetLastError();
      dev = 0;
    }
    indigo_ctx* c = new indigo_ctx();
    const int r = ctx_init(dev, c);
    if (r != INDIGO_OK) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      *rc = r;
      return nullptr;
    }
    g_default = c;
  }
  *rc = INDIGO_OK;
  return g_default;
}

}  // namespace indigo

using namespace indigo;

extern "C" {

const char* indigo_last_error(void) { return error_buffer(); }

int indigo_version(void) { return 100; }

int indigo_device_info(int device, int* compute_units, int* max_threads_per_cu, char* name, int name_len)
{
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    return set_error(INDIGO_ERR_NO_DEVICE, "no HIP device available");
  }
  if (device < 0 || device >= count) return set_error(INDIGO_ERR_INVALID_ARG, "device %d out of range", device);
  hipDeviceProp_t prop;
  IG_HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (compute_units) *compute_units = prop.multiProcessorCount;
  if (max_threads_per_cu) *max_threads_per_cu = prop.maxThreadsPerMultiProcessor;
  if (name && name_len > 0) {
    strncpy(name, prop.name, static_cast<size_t>(name_len) - 1);
    name[name_len - 1] = '\0';
  }
  return INDIGO_OK;
}

int indigo_malloc(void** d_ptr, size_t bytes)
{
  if (d_ptr == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "indigo_malloc: null out pointer");
  if (hipMalloc(d_ptr, std::max<size_t>(bytes, 4)) != hipSuccess) {
    (void)hipGetLastError();
    *d_ptr = nullptr;
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed", bytes);
  }
  return INDIGO_OK;
}

int indigo_free(void* d_ptr)
{
  if (d_ptr == nullptr) return INDIGO_OK;
  IG_HIP_TRY(hipFree(d_ptr));
  return INDIGO_OK;
}

int indigo_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes)
{
  if (bytes == 0) return INDIGO_OK;
  if (!d_dst || !h_src) return set_error(INDIGO_ERR_INVALID_ARG, "memcpy_h2d: null pointer");
  IG_HIP_TRY(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
  return INDIGO_OK;
}

int indigo_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes)
{
  if (bytes == 0) return INDIGO_OK;
  if (!h_dst || !d_src) return set_error(INDIGO_ERR_INVALID_ARG, "memcpy_d2h: null pointer");
  IG_HIP_TRY(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
  return INDIGO_OK;
}

int indigo_ctx_create(int device, indigo_ctx** out)
{
  if (out == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_create: null out pointer");
  *out = nullptr;
  indigo_ctx* c = new indigo_ctx();
  const int rc = ctx_init(device, c);
  if (rc != INDIGO_OK) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return INDIGO_OK;
}

int indigo_ctx_destroy(indigo_ctx* ctx)
{
  if (ctx == nullptr) return INDIGO_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  drop_captured(ctx);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  for (int b = 0; b < 2; b++) {
    if (ctx->stage[b]) (void)hipHostFree(ctx->stage[b]);
    if (ctx->ev_stage[b]) (void)hipEventDestroy(ctx->ev_stage[b]);
  }
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->ev_k0) (void)hipEventDestroy(ctx->ev_k0);
  if (ctx->ev_k1) (void)hipEventDestroy(ctx->ev_k1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return INDIGO_OK;
}

int indigo_shutdown(void)
{
  indigo_ctx* c = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_default_mu);
    c = g_default;
    g_default = nullptr;
  }
  if (c != nullptr) (void)indigo_ctx_destroy(c);
  int count = 0;
  if (hipGetDeviceCount(&count) == hipSuccess) {
    for (int d = 0; d < count; d++)
      if (hipSetDevice(d) == hipSuccess) (void)hipDeviceSynchronize();
  }
  (void)hipGetLastError();
  return INDIGO_OK;
}

int indigo_ctx_last_ms(indigo_ctx* ctx, float* ms)
{
  if (ctx == nullptr || ms == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_last_ms: null argument");
  *ms = ctx->last_ms;
  return INDIGO_OK;
}

int indigo_ctx_last_stats(indigo_ctx* ctx, int* launches, int* iterations)
{
  if (ctx == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_last_stats: null ctx");
  if (launches) *launches = ctx->last_launches;
  if (iterations) *iterations = ctx->last_iterations;
  return INDIGO_OK;
}

int indigo_ctx_last_kernel(indigo_ctx* ctx, float* ms, char* name, int name_len)
{
  if (ctx == nullptr || ms == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_last_kernel: null argument");
  *ms = ctx->last_kernel_ms;
  if (name != nullptr && name_len > 0) {
    std::snprintf(name, static_cast<size_t>(name_len), "%s", ctx->last_kernel);
  }
  return INDIGO_OK;
}

int indigo_ctx_copy_ms(indigo_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, float* ms)
{
  if (ctx == nullptr || ms == nullptr || (bytes > 0 && (d_dst == nullptr || d_src == nullptr)))
    return set_error(INDIGO_ERR_INVALID_ARG, "ctx_copy_ms: null argument");
  IG_HIP_TRY(hipSetDevice(ctx->device));
  IG_HIP_TRY(hipEventRecord(ctx->ev_k0, ctx->stream));
  IG_HIP_TRY(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  IG_HIP_TRY(hipEventRecord(ctx->ev_k1, ctx->stream));
  IG_HIP_TRY(hipEventSynchronize(ctx->ev_k1));
  IG_HIP_TRY(hipEventElapsedTime(ms, ctx->ev_k0, ctx->ev_k1));
  return INDIGO_OK;
}

int indigo_ctx_last_grid(indigo_ctx* ctx, int* grid, int* per_cu)
{
  if (ctx == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_last_grid: null ctx");
  if (grid) *grid = ctx->last_grid;
  if (per_cu) *per_cu = ctx->last_per_cu;
  return INDIGO_OK;
}

int indigo_ctx_last_trace(indigo_ctx* ctx, int max, float* us, unsigned* frontier, double* degree_sum, int* mode,
                          int* n)
{
  if (ctx == nullptr || n == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_last_trace: null argument");
  const int k = ctx->trace_n < max ? ctx->trace_n : max;
  for (int i = 0; i < k; i++) {
    if (us) us[i] = ctx->trace_us[i];
    if (frontier) frontier[i] = ctx->trace_nf[i];
    if (degree_sum) degree_sum[i] = ctx->trace_m[i];
    if (mode) mode[i] = ctx->trace_mode[i];
  }
  *n = k;
  return INDIGO_OK;
}

int indigo_ctx_trim(indigo_ctx* ctx)
{
  if (ctx == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_trim: null ctx");
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  drop_captured(ctx);
  if (ctx->ws) IG_HIP_TRY(hipFree(ctx->ws));
  ctx->ws = nullptr;
  ctx->ws_bytes = 0;
  return INDIGO_OK;
}

int indigo_ctx_set_hint(indigo_ctx* ctx, int key, int value)
{
  if (ctx == nullptr || key < 0 || key >= 16) return set_error(INDIGO_ERR_INVALID_ARG, "ctx_set_hint: bad key");
  ctx->hints[key] = value;
  return INDIGO_OK;
}

}  // extern "C"
