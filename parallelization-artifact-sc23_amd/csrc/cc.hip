// Connected components (label = smallest vertex id of the component) on gfx950.
//
// Reference semantics: GPUcc_vertex / GPUcc_edge (include/cc_vertex_cuda.h:57,
// include/cc_edge_cuda.h:57): label[v] = v, then min-label propagation to a
// fixed point (e.g. codes/cuda/cc-cuda/CC_CUDA_V_Topo_Push_NonDeterm_IntType_
// ReadModifyWrite_NonPersist_Atomic_Thread_NoBoundsBug2_NoBoundsBug1.cu:47-76),
// whose result on a symmetric graph is the minimum id of each component.
//
// Design: instead of diameter-many propagation sweeps, a concurrent
// union-find whose every link hooks the LARGER root under the smaller one
// with a CAS (parent[x] <= x always holds), so each final root is the
// minimum id of its tree and, once all edges are linked, of its component -
// the same labels, bit for bit.  Edges are linked in the order of Afforest
// (Sutton, Ben-Nun, Barak, IPDPS'18), which lets most of them be skipped:
//   1. init:     parent[v] = min(v, smallest neighbour) (lists are sorted);
//   2. sample:   link v with its first kSample neighbours;
//   3. compress: parent[v] = root(v);
//   4. giant:    the most frequent root c among 1024 sampled vertices;
//   5. link:     every vertex whose root is not c links ALL its remaining
//                entries (short lists with the in-wave balancer, longer ones
//                binned into mid / huge queues, appends aggregated per
//                workgroup); vertices in c's tree skip their lists;
//   6. heavy:    mid lists one wave each, huge lists in kHugeChunk-entry
//                items, one workgroup each;
//   7. flatten:  label[v] = root(v).
// Graphs with an average degree below kAfforestDeg skip steps 2-4 and link
// every undirected edge once, from its larger endpoint.
// Skipping is exact on a symmetric graph: an edge (a, b) is linked from a
// unless a was connected to c when it checked; it is then linked from b
// unless b was too - and then both are in c's component already.  On R-MAT
// graphs the giant component holds nearly every vertex with edges, so step
// 5 reads only a few lists.
// Finds use path halving with relaxed stores: every store replaces a parent
// by one of its ancestors, so concurrent finds and links stay valid.
#include "cc.hpp"

#include <climits>
#include <mutex>

#include "frontier.hpp"
#include "unionfind.hpp"

namespace indigo {

namespace {

constexpr int kCcBlock = 256;

constexpr int kSample = 2;         // neighbours linked per vertex before the giant root is chosen
constexpr int kGiantSamples = 1024;  // vertices sampled to choose it
constexpr long long kAfforestDeg = 8;  // average degree from which the Afforest order is used

struct CcControl {
  FrontierCounters heavy;  // deferred long lists (mid / huge queues)
  int giant;               // root of the most frequent component in the sample
  int pad[31];
};

__global__ __launch_bounds__(kCcBlock) void cc_init(int nodes, const int* __restrict__ nindex,
                                                    const int* __restrict__ nlist, int* parent)
{
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nodes; v += gridDim.x * blockDim.x) {
    const int beg = nindex[v];
    int p = v;
    if (beg < nindex[v + 1]) {
      const int u = nlist[beg];
      if (u < v) p = u;
    }
    parent[v] = p;
  }
}

__global__ __launch_bounds__(kCcBlock) void cc_sample(int nodes, const int* __restrict__ nindex,
                                                      const int* __restrict__ nlist, int* parent)
{
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nodes; v += gridDim.x * blockDim.x) {
    const int beg = nindex[v], end = min(nindex[v + 1], nindex[v] + kSample);
    for (int e = beg; e < end; e++) link(parent, v, nlist[e]);
  }
}

__global__ __launch_bounds__(kCcBlock) void cc_compress(int nodes, int* parent)
{
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nodes; v += gridDim.x * blockDim.x)
    st_relaxed(&parent[v], find_root(parent, v));
}

// One workgroup: the most frequent root among kGiantSamples vertices spread
// over the id range (ties to the smaller root).
__global__ __launch_bounds__(kCcBlock) void cc_giant(int nodes, int* parent, CcControl* ctl)
{
  __shared__ int s_root[kGiantSamples];
  __shared__ int s_best[kCcBlock], s_cnt[kCcBlock];
  for (int i = threadIdx.x; i < kGiantSamples; i += blockDim.x) {
    const int v = static_cast<int>((static_cast<unsigned long long>(i) * 2654435761ull) % static_cast<unsigned>(nodes));
    s_root[i] = find_root(parent, v);
  }
  __syncthreads();
  int best = INT_MAX, bcnt = 0;
  for (int i = threadIdx.x; i < kGiantSamples; i += blockDim.x) {
    int c = 0;
    for (int j = 0; j < kGiantSamples; j++) c += s_root[j] == s_root[i];
    if (c > bcnt || (c == bcnt && s_root[i] < best)) {
      bcnt = c;
      best = s_root[i];
    }
  }
  s_best[threadIdx.x] = best;
  s_cnt[threadIdx.x] = bcnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < static_cast<int>(blockDim.x); t++)
      if (s_cnt[t] > bcnt || (s_cnt[t] == bcnt && s_best[t] < best)) {
        bcnt = s_cnt[t];
        best = s_best[t];
      }
    st_relaxed(&ctl->giant, best);
  }
}

// kSkip (Afforest order): vertices in the giant root's tree skip their lists,
// the others link all entries after the sampled ones.  Otherwise every
// undirected edge is linked once, from its larger endpoint.
template <bool kSkip>
__global__ __launch_bounds__(kCcBlock) void cc_link(int nodes, const int* __restrict__ nindex,
                                                    const int* __restrict__ nlist, int* parent, FrontierQueues heavy,
                                                    CcControl* ctl)
{
  __shared__ FrontierEmitter s_emit;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nchunks = (nodes + kWave - 1) / kWave;
  const int giant = kSkip ? ld_relaxed(&ctl->giant) : -1;
  const int first = kSkip ? kSample : 0;  // entries already linked
  fe_reset(s_emit);
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < nchunks; c += gridDim.x * wpb) {
    const int v = c * kWave + lane;
    const bool valid = v < nodes;
    const int beg = valid ? nindex[v] : 0;
    const int deg = valid ? nindex[v + 1] - beg : 0;
    // lists of vertices already joined to the giant root are skipped
    const bool todo = deg > first && (!kSkip || find_root(parent, v) != giant);
    const bool long_list = todo && deg >= kMidDeg;
    frontier_emit_block(s_emit, long_list, v, deg, heavy, &ctl->heavy);
    const int rest = todo && !long_list ? deg - first : 0;
    const int incl = wave_inclusive_scan(rest);
    const int total = wave_shfl(incl, kWave - 1);
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int obeg = wave_shfl(beg, o) + first;
      const int oexcl = wave_shfl(incl - rest, o);
      const int ov = c * kWave + o;
      if (k < total) {
        const int u = nlist[obeg + (k - oexcl)];
        if (kSkip || u < ov) link(parent, ov, u);
      }
    }
  }
  fe_flush(s_emit, heavy, &ctl->heavy);
}

template <bool kSkip>
__global__ __launch_bounds__(kCcBlock) void cc_link_heavy(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int* parent,
                                                          FrontierQueues heavy, const CcControl* ctl)
{
  const unsigned long long hp = ctl->heavy.hpack;
  auto visit = [&](int v, int e, bool active) {
    if (active) {
      const int u = nlist[e];
      if (kSkip || u < v) link(parent, v, u);
    }
  };
  expand_huge(heavy, huge_count(hp), huge_chunks(hp), nindex, visit);
  expand_mid(heavy, ctl->heavy.nm, nindex, visit);
}

__global__ __launch_bounds__(kCcBlock) void cc_flatten(int nodes, int* parent, int* label)
{
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nodes; v += gridDim.x * blockDim.x)
    label[v] = find_root(parent, v);
}

int grid_for(long long items, int per_block, int cap = 1 << 20)
{
  const long long g = (items + per_block - 1) / per_block;
  return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

size_t cc_workspace_bytes(int nodes)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(CcControl)) + 4 * Carver::round(v * 4) + 4096;
}

int cc_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* label, int* launches)
{
  return cc_run_in(ctx, ctx->ws, nodes, edges, nindex, nlist, label, launches);
}

int cc_run_in(indigo_ctx* ctx, void* ws, int nodes, int edges, const int* nindex, const int* nlist, int* label,
              int* launches)
{
  Carver cv(ws);
  CcControl* ctl = cv.take<CcControl>(1);
  int* parent = cv.take<int>(nodes);
  FrontierQueues heavy{nullptr, cv.take<int>(nodes), cv.take<int>(nodes), cv.take<unsigned>(nodes)};
  IG_TRY(zero_async(ctx, ctl, sizeof(CcControl)));
  const int gv = share_grid(ctx, grid_for(nodes, kCcBlock, 8192), kCcBlock);
  const int gl = share_grid(ctx, grid_for((nodes + kWave - 1) / kWave, kCcBlock / kWave, 1 << 16), kCcBlock);
  const int gh = share_grid(ctx, ctx->compute_units * 8, kCcBlock);
  // The Afforest order pays an extra compress and a find per vertex to skip
  // the giant component's lists: worth it when lists are long (R-MAT: 1.04
  // -> 0.25 ms at scale 22), not on sparse high-diameter graphs, whose deep
  // trees make those finds expensive (4096^2 road-like grid: 5.5 -> 8.5 ms).
  const bool afforest = static_cast<long long>(edges) >= kAfforestDeg * static_cast<long long>(nodes);
  // no single kernel dominates (the Afforest order skips most lists), so
  // the timed span is the whole kernel sequence, init to flatten
  IG_TRY(kernel_begin(ctx, "cc_init..cc_flatten"));
  cc_init<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent);
  if (afforest) {
    cc_sample<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent);
    cc_compress<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, parent);
    cc_giant<<<1, kCcBlock, 0, ctx->stream>>>(nodes, parent, ctl);
    cc_link<true><<<gl, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent, heavy, ctl);
    cc_link_heavy<true><<<gh, kCcBlock, 0, ctx->stream>>>(nindex, nlist, parent, heavy, ctl);
  } else {
    cc_link<false><<<gl, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent, heavy, ctl);
    cc_link_heavy<false><<<gh, kCcBlock, 0, ctx->stream>>>(nindex, nlist, parent, heavy, ctl);
  }
  cc_flatten<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, parent, label);
  IG_TRY(kernel_end(ctx));
  IG_HIP_TRY(hipGetLastError());
  if (launches) *launches = afforest ? 8 : 5;
  return INDIGO_OK;
}

// ---- scratch of the allocation-free entry (indigo_cc_noalloc) ------------------
// A file-scope device array stands in for the context workspace, so a call
// makes no allocation.  Each device holds its own copy for the life of the
// process; host code must use the address hipGetSymbolAddress returns for the
// current device, never the array's name (a host-side stand-in).  Every call
// on a device shares that device's copy: cc_run_in clears what must start at
// zero (the control block) with a kernel and initialises the rest, and two
// calls that overlap on one device race.
namespace {

constexpr size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
constexpr size_t kScratchBytes =
    round256(sizeof(CcControl)) + 4 * round256(static_cast<size_t>(kCcScratchMaxNodes) * 4) + 4096;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on ctx's device, looked up once per device
// with that device current.
int scratch_address(indigo_ctx* ctx, void** out)
{
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(INDIGO_ERR_INVALID_ARG, "cc_noalloc: device %d out of range", ctx->device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[ctx->device] == nullptr) {
    IG_HIP_TRY(hipSetDevice(ctx->device));
    void* p = nullptr;
    IG_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (p == nullptr) return set_error(INDIGO_ERR_DEVICE, "cc_noalloc: scratch has no device address");
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return INDIGO_OK;
}

}  // namespace

int cc_noalloc_device(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* label)
{
  if (nodes > kCcScratchMaxNodes)
    return set_error(INDIGO_ERR_CAPACITY, "cc_noalloc: %d vertices exceed the scratch limit of %d", nodes,
                     kCcScratchMaxNodes);
  if (cc_workspace_bytes(nodes) > kScratchBytes)
    return set_error(INDIGO_ERR_CAPACITY, "cc_noalloc: workspace of %zu bytes exceeds the %zu-byte scratch",
                     cc_workspace_bytes(nodes), kScratchBytes);
  void* ws = nullptr;
  IG_TRY(scratch_address(ctx, &ws));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  IG_TRY(timing_begin(ctx));
  int launches = 0;
  IG_TRY(cc_run_in(ctx, ws, nodes, edges, nindex, nlist, label, &launches));
  return timing_end(ctx, launches, 1);
}

}  // namespace indigo
