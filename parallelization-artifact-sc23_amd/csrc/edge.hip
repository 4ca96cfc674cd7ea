// Edge-centric, topology-driven variants of BFS, SSSP, CC and MIS on gfx950:
// the suite's E_Topo_Push family, which walks the CSR as an edge list
// (sp[e], nlist[e]) instead of per vertex.
//
// Reference semantics: GPUbfs_edge / GPUsssp_edge / GPUcc_edge / GPUmis_edge
// (include/bfs_edge_cuda.h:61, include/sssp_edge_cuda.h:60,
// include/cc_edge_cuda.h:57, include/mis_edge_cuda.h:54) take the edge-source
// array sp (sp[e] = v for e in [nindex[v], nindex[v + 1]), built on the host
// by the harness, bfs_edge_cuda.h:140-146) and assign one thread per CSR
// entry, e.g. codes/cuda/cc-cuda/CC_CUDA_E_Topo_Push_*.cu.  The outputs are
// the vertex forms' outputs, bit for bit.
//
// Design (own): one persistent kernel per call, each workgroup striding the
// entry range [0, E) with 64-lane waves, so every wave reads 64 consecutive
// sp / nlist (/ eweight) words - perfect balance whatever the degree skew,
// at the price of visiting every entry in every sweep.  Sweeps are separated
// by the coherent grid barrier; all data that workgroups share is accessed
// with device-scope relaxed operations.  A sweep that changes nothing ends
// the loop: its flag is triple-buffered (set in sweep i, read after barrier
// i, cleared in sweep i + 1 by workgroup 0), so one barrier per sweep suffices.
//   BFS : level-synchronous; sweep L sets dist[v] = L + 1 for every entry
//         (u, v) with dist[u] == L and dist[v] unreached.
//   SSSP: Bellman-Ford sweeps in place, atomicMin(dist[v], dist[u] + w).
//   CC  : one linking sweep over the entries with u > v (each undirected
//         edge once) with the union-find of unionfind.hpp, then a flatten.
//   MIS : rounds of four phases: every undecided vertex is a candidate; an
//         entry (u, v) with both undecided and key(v) > key(u) withdraws u;
//         candidates join; an entry (u, v) with u included excludes v.  With
//         fixed priorities this is the greedy set in descending key order.
// The workgroup size is a template parameter (256, 512 or 1024 threads;
// INDIGO_HINT_BLOCK_SIZE), one of the autotuner's axes.
#include "edge.hpp"

#include "graphio.hpp"
#include "unionfind.hpp"

namespace indigo {

namespace {

struct alignas(128) EdgeControl {
  GridBarrier bar;
  PaddedU32 changed[3];
  PaddedU32 sweeps;
};

struct EdgeArgs {
  int nodes;
  unsigned edges;
  const int* sp;
  const int* nlist;
  const int* eweight;
  int src;
  unsigned flip;  // MIS: 0 unsigned priority order, 0x80000000 signed
  int* out;       // dist / label / status
  int* aux;       // CC parent, MIS candidate flags
  EdgeControl* ctl;
  long long budget;
  long long res_ticks;
};

constexpr int kUndecided = 0, kIncluded = 1, kExcluded = 2;  // include/mis_edge_cuda.h:50-52

IG_DEV unsigned long long edge_mis_key(int v, unsigned flip)
{
  const unsigned p = mis_hash(static_cast<unsigned>(v) + 712313887u) ^ flip;
  return (static_cast<unsigned long long>(p) << 32) | static_cast<unsigned>(v);
}

// Raises the sweep's flag when any thread of the workgroup saw a change.
template <int kBlock>
IG_DEV void flag_sweep(EdgeControl* ctl, int sweep, bool any)
{
  if (__syncthreads_or(any) && threadIdx.x == 0) st_relaxed(&ctl->changed[sweep % 3].v, 1u);
}

// After the barrier that closed `sweep`: true when the sweep changed nothing.
IG_DEV bool sweep_quiet(EdgeControl* ctl, int sweep)
{
  return ld_relaxed(&ctl->changed[sweep % 3].v) == 0u;
}

IG_DEV void open_sweep(EdgeControl* ctl, int sweep)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&ctl->changed[(sweep + 1) % 3].v, 0u);
}

IG_DEV void record_sweeps(EdgeControl* ctl, int n)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&ctl->sweeps.v, static_cast<unsigned>(n));
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void edge_bfs(EdgeArgs a)
{
  if (!residency_barrier(&a.ctl->bar, a.res_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const unsigned tid = blockIdx.x * kBlock + threadIdx.x, nth = gridDim.x * kBlock;
  for (unsigned v = tid; v < static_cast<unsigned>(a.nodes); v += nth)
    st_relaxed(&a.out[v], static_cast<int>(v) == a.src ? 0 : kInf);
  if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  for (int level = 0;; level++) {
    open_sweep(a.ctl, level);
    bool any = false;
    for (unsigned e = tid; e < a.edges; e += nth) {
      if (ld_relaxed(&a.out[a.sp[e]]) != level) continue;
      const int v = a.nlist[e];
      if (ld_relaxed(&a.out[v]) == kInf) {
        st_relaxed(&a.out[v], level + 1);  // every writer stores the same level
        any = true;
      }
    }
    flag_sweep<kBlock>(a.ctl, level, any);
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
    if (sweep_quiet(a.ctl, level)) {
      record_sweeps(a.ctl, level + 1);
      return;
    }
  }
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void edge_sssp(EdgeArgs a)
{
  if (!residency_barrier(&a.ctl->bar, a.res_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const unsigned tid = blockIdx.x * kBlock + threadIdx.x, nth = gridDim.x * kBlock;
  for (unsigned v = tid; v < static_cast<unsigned>(a.nodes); v += nth)
    st_relaxed(&a.out[v], static_cast<int>(v) == a.src ? 0 : kInf);
  if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  for (int sweep = 0;; sweep++) {
    open_sweep(a.ctl, sweep);
    bool any = false;
    for (unsigned e = tid; e < a.edges; e += nth) {
      const int du = ld_relaxed(&a.out[a.sp[e]]);
      if (du == kInf) continue;
      const int nd = du + a.eweight[e];  // int arithmetic, as the reference relaxes
      const int v = a.nlist[e];
      if (nd < ld_relaxed(&a.out[v]) && atomic_min_relaxed(&a.out[v], nd) > nd) any = true;
    }
    flag_sweep<kBlock>(a.ctl, sweep, any);
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
    if (sweep_quiet(a.ctl, sweep)) {
      record_sweeps(a.ctl, sweep + 1);
      return;
    }
  }
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void edge_cc(EdgeArgs a)
{
  if (!residency_barrier(&a.ctl->bar, a.res_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const unsigned tid = blockIdx.x * kBlock + threadIdx.x, nth = gridDim.x * kBlock;
  int* parent = a.aux;
  for (unsigned v = tid; v < static_cast<unsigned>(a.nodes); v += nth) st_relaxed(&parent[v], static_cast<int>(v));
  if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  for (unsigned e = tid; e < a.edges; e += nth) {
    const int u = a.sp[e], v = a.nlist[e];
    if (u > v) link(parent, u, v);
  }
  if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  for (unsigned v = tid; v < static_cast<unsigned>(a.nodes); v += nth) a.out[v] = find_root(parent, static_cast<int>(v));
  record_sweeps(a.ctl, 1);
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void edge_mis(EdgeArgs a)
{
  if (!residency_barrier(&a.ctl->bar, a.res_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const unsigned tid = blockIdx.x * kBlock + threadIdx.x, nth = gridDim.x * kBlock;
  const unsigned nodes = static_cast<unsigned>(a.nodes);
  int* status = a.out;
  int* cand = a.aux;
  for (unsigned v = tid; v < nodes; v += nth) st_relaxed(&status[v], kUndecided);
  if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  for (int round = 0;; round++) {
    // every undecided vertex is a candidate; the round's flag says one exists
    open_sweep(a.ctl, round);
    bool any = false;
    for (unsigned v = tid; v < nodes; v += nth) {
      const bool und = ld_relaxed(&status[v]) == kUndecided;
      st_relaxed(&cand[v], und ? 1 : 0);
      any |= und;
    }
    flag_sweep<kBlock>(a.ctl, round, any);
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
    if (sweep_quiet(a.ctl, round)) {
      record_sweeps(a.ctl, round);
      return;
    }
    // an undecided neighbour with a larger key withdraws the candidate
    for (unsigned e = tid; e < a.edges; e += nth) {
      const int u = a.sp[e];
      if (ld_relaxed(&cand[u]) == 0) continue;
      const int v = a.nlist[e];
      if (ld_relaxed(&status[v]) == kUndecided && edge_mis_key(v, a.flip) > edge_mis_key(u, a.flip))
        st_relaxed(&cand[u], 0);
    }
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
    for (unsigned v = tid; v < nodes; v += nth)
      if (ld_relaxed(&cand[v]) != 0) st_relaxed(&status[v], kIncluded);
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
    // neighbours of the new members leave
    for (unsigned e = tid; e < a.edges; e += nth) {
      const int u = a.sp[e];
      if (ld_relaxed(&status[u]) != kIncluded) continue;
      const int v = a.nlist[e];
      if (ld_relaxed(&status[v]) == kUndecided) st_relaxed(&status[v], kExcluded);
    }
    if (!grid_barrier<true>(&a.ctl->bar, deadline)) return;
  }
}

using EdgeKernel = void (*)(EdgeArgs);

template <int kBlock>
EdgeKernel kernel_for(int alg)
{
  switch (alg) {
    case kEdgeBfs: return edge_bfs<kBlock>;
    case kEdgeSssp: return edge_sssp<kBlock>;
    case kEdgeCc: return edge_cc<kBlock>;
    default: return edge_mis<kBlock>;
  }
}

const char* kernel_name(int alg)
{
  switch (alg) {
    case kEdgeBfs: return "edge_bfs";
    case kEdgeSssp: return "edge_sssp";
    case kEdgeCc: return "edge_cc";
    default: return "edge_mis";
  }
}

}  // namespace

int edge_block(const indigo_ctx* ctx)
{
  const int b = ctx->hints[INDIGO_HINT_BLOCK_SIZE];
  return (b == 256 || b == 1024) ? b : 512;
}

size_t edge_workspace_bytes(int nodes, int edges, bool build_sp)
{
  return Carver::round(sizeof(EdgeControl)) + Carver::round(static_cast<size_t>(nodes) * 4) +
         (build_sp ? Carver::round(static_cast<size_t>(edges) * 4) : 0) + 4096;
}

int edge_device(indigo_ctx* ctx, int alg, int nodes, int edges, const int* nindex, const int* nlist,
                const int* eweight, const int* sp, int src, int signed_priority, int* out, double* seconds)
{
  if (alg < kEdgeBfs || alg > kEdgeMis) return set_error(INDIGO_ERR_INVALID_ARG, "edge: unknown algorithm %d", alg);
  IG_TRY(ensure_workspace(ctx, edge_workspace_bytes(nodes, edges, sp == nullptr)));
  IG_TRY(ensure_host_pinned(ctx, 4096));
  Carver cv(ctx->ws);
  EdgeControl* ctl = cv.take<EdgeControl>(1);
  int* aux = cv.take<int>(nodes);
  int* own_sp = sp == nullptr ? cv.take<int>(edges) : nullptr;
  const int block = edge_block(ctx);
  EdgeKernel k = block == 256 ? kernel_for<256>(alg) : (block == 1024 ? kernel_for<1024>(alg) : kernel_for<512>(alg));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const double t0 = wall_seconds();
  unsigned sweeps = 0;
  for (;;) {
    IG_TRY(timing_begin(ctx));
    if (own_sp != nullptr) IG_TRY(build_sp_async(ctx, nodes, edges, nindex, own_sp));
    IG_TRY(zero_async(ctx, ctl, sizeof(EdgeControl)));
    EdgeArgs a{nodes, static_cast<unsigned>(edges), own_sp != nullptr ? own_sp : sp, nlist, eweight, src,
               signed_priority ? 0x80000000u : 0u, out, aux, ctl, budget_ticks(ctx), residency_ticks(ctx)};
    void* args[] = {&a};
    IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(k), block, 0, args, 0, nullptr, kernel_name(alg)));
    IG_HIP_TRY(hipMemcpyAsync(ctx->host_pinned, &ctl->bar.abort.v, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
    IG_HIP_TRY(hipMemcpyAsync(static_cast<unsigned*>(ctx->host_pinned) + 1, &ctl->sweeps.v, sizeof(unsigned),
                              hipMemcpyDeviceToHost, ctx->stream));
    IG_TRY(timing_end(ctx, own_sp != nullptr ? 3 : 2, 0));
    const unsigned* h = static_cast<const unsigned*>(ctx->host_pinned);
    const int rc = check_abort(ctx, reinterpret_cast<const void*>(k), h[0], kernel_name(alg));
    if (rc == INDIGO_OK) {
      sweeps = h[1];
      break;
    }
    if (rc != kNotResident) return rc;
  }
  if (seconds) *seconds = wall_seconds() - t0;
  ctx->last_iterations = static_cast<int>(sweeps);
  return INDIGO_OK;
}

}  // namespace indigo
