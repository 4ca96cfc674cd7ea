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
// the same labels, bit for bit.
//   1. init:    parent[v] = min(v, smallest neighbour) (lists are sorted);
//   2. link:    every undirected edge once (from its larger endpoint); a wave
//               takes 64 vertices in id order and walks the CSR entries of
//               their short lists (< kMidDeg) 64 at a time with the in-wave
//               balancer; longer lists go to the mid / huge queues
//               (appends aggregated per workgroup in LDS);
//   3. heavy:   mid lists one wave each, huge lists in kHugeChunk-entry
//               items, one workgroup each;
//   4. flatten: label[v] = root(v).
// Finds use path halving with relaxed stores: every store replaces a parent
// by one of its ancestors, so concurrent finds and links stay valid.
#include "cc.hpp"

#include "frontier.hpp"

namespace indigo {

namespace {

constexpr int kCcBlock = 256;

struct CcControl {
  FrontierCounters heavy;  // deferred long lists (mid / huge queues)
};

IG_DEV int find_root(int* parent, int x)
{
  int p = ld_relaxed(&parent[x]);
  while (p != x) {
    const int gp = ld_relaxed(&parent[p]);
    if (gp == p) return p;
    st_relaxed(&parent[x], gp);  // path halving: gp is an ancestor of x
    x = gp;
    p = ld_relaxed(&parent[x]);
  }
  return x;
}

IG_DEV void link(int* parent, int a, int b)
{
  int ra = find_root(parent, a);
  int rb = find_root(parent, b);
  while (ra != rb) {
    const int hi = ra > rb ? ra : rb;
    const int lo = ra > rb ? rb : ra;
    int expected = hi;
    if (__hip_atomic_compare_exchange_strong(&parent[hi], &expected, lo, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
    // hi stopped being a root meanwhile: climb from where it now points
    ra = find_root(parent, ra == hi ? expected : ra);
    rb = find_root(parent, rb == hi ? expected : rb);
  }
}

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

__global__ __launch_bounds__(kCcBlock) void cc_link(int nodes, const int* __restrict__ nindex,
                                                    const int* __restrict__ nlist, int* parent, FrontierQueues heavy,
                                                    CcControl* ctl)
{
  __shared__ FrontierEmitter s_emit;
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nchunks = (nodes + kWave - 1) / kWave;
  fe_reset(s_emit);
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < nchunks; c += gridDim.x * wpb) {
    const int v = c * kWave + lane;
    const bool valid = v < nodes;
    const int beg = valid ? nindex[v] : 0;
    int deg = valid ? nindex[v + 1] - beg : 0;
    const bool long_list = deg >= kMidDeg;
    frontier_emit_block(s_emit, long_list, v, deg, heavy, &ctl->heavy);
    if (long_list) deg = 0;
    const int incl = wave_inclusive_scan(deg);
    const int total = wave_shfl(incl, kWave - 1);
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int obeg = wave_shfl(beg, o);
      const int oexcl = wave_shfl(incl - deg, o);
      const int ov = c * kWave + o;
      if (k < total) {
        const int u = nlist[obeg + (k - oexcl)];
        if (u < ov) link(parent, ov, u);
      }
    }
  }
  fe_flush(s_emit, heavy, &ctl->heavy);
}

__global__ __launch_bounds__(kCcBlock) void cc_link_heavy(const int* __restrict__ nindex,
                                                          const int* __restrict__ nlist, int* parent,
                                                          FrontierQueues heavy, const CcControl* ctl)
{
  const unsigned long long hp = ctl->heavy.hpack;
  auto visit = [&](int v, int e, bool active) {
    if (active) {
      const int u = nlist[e];
      if (u < v) link(parent, v, u);
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
  (void)edges;
  Carver cv(ctx->ws);
  CcControl* ctl = cv.take<CcControl>(1);
  int* parent = cv.take<int>(nodes);
  FrontierQueues heavy{nullptr, cv.take<int>(nodes), cv.take<int>(nodes), cv.take<unsigned>(nodes)};
  IG_TRY(zero_async(ctx, ctl, sizeof(CcControl)));
  const int gv = grid_for(nodes, kCcBlock, 8192);
  cc_init<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent);
  const int gl = grid_for((nodes + kWave - 1) / kWave, kCcBlock / kWave, 1 << 16);
  IG_TRY(kernel_begin(ctx, "cc_link"));
  cc_link<<<gl, kCcBlock, 0, ctx->stream>>>(nodes, nindex, nlist, parent, heavy, ctl);
  IG_TRY(kernel_end(ctx));
  cc_link_heavy<<<ctx->compute_units * 8, kCcBlock, 0, ctx->stream>>>(nindex, nlist, parent, heavy, ctl);
  cc_flatten<<<gv, kCcBlock, 0, ctx->stream>>>(nodes, parent, label);
  IG_HIP_TRY(hipGetLastError());
  if (launches) *launches = 5;
  return INDIGO_OK;
}

}  // namespace indigo
