// Maximal independent set by hash priority on gfx950.
//
// Reference semantics: GPUmis_vertex / GPUmis_edge (include/mis_vertex_cuda.h:54,
// include/mis_edge_cuda.h:54): priority[v] = hash(v + 712313887); an undecided
// vertex joins the set when every neighbour is excluded or has a smaller
// (priority, id), and then excludes its neighbours
// (codes/cuda/mis-cuda/MIS_CUDA_V_Topo_Push_Determ_IntType_NonPersist_Atomic_Thread.cu:57-81).
// The fixed point is the greedy MIS in descending (priority, id) order; the
// suite's C++ path orders priorities as unsigned int, its CUDA path as signed
// int (data_type), selected here by `signed_priority`.
//
// Design: one persistent kernel (coherent-mode barriers: status,
// worklists and counters only see device-scope atomic accesses); round r
// processes the vertices
// still undecided after round r-1, kept in two worklists by list length:
// lists shorter than kMidDeg (round 0: every such vertex, in id order) are
// scanned by one lane each, longer ones (round 0: a static list collected at
// start) by one wave each, 256 entries per step with ballots.  Vertices that
// stay undecided are appended (wave-aggregated) to the next round's lists.
// Decisions are made in place: a decided state (1 included, 2 excluded) never
// changes, so a vertex that reads a neighbour's final state may act on it in
// the same round and a stale "undecided" read only delays it.  A vertex is
// included only when every neighbour with a larger key is excluded, and
// excluded as soon as a neighbour is included: exactly the greedy order's
// fixed point.  Keys are (priority ^ sign flip, id) compared as one 64-bit
// unsigned value; the priority is recomputed from the hash instead of being
// stored.  Self loops are ignored.
#include "mis.hpp"

#include "frontier.hpp"
#include "trace.hpp"

namespace indigo {

namespace {

constexpr int kMisBlock = 512;

constexpr int kUndecided = 0, kIncluded = 1, kExcluded = 2;
constexpr int kMisLaneUnroll = 4;  // entries per step of a one-lane decision

struct MisControl {
  GridBarrier bar;
  FrontierCounters cnt[3];  // n: short-list worklist, nm: long-list worklist
  int rounds;
  int status;
  int pad[30];
  LevelTrace trace;  // per round: start, undecided short-list and long-list vertices
};

struct MisArgs {
  int nodes;
  unsigned flip;
  const int* __restrict__ nindex;
  const int* __restrict__ nlist;
  int* status;
  int* wl[2];   // short-list worklists
  int* hwl[2];  // long-list worklists (round 0: the static list)
  MisControl* ctl;
  long long budget;
  long long start_ticks;
};

IG_DEV unsigned long long mis_key(int v, unsigned flip)
{
  const unsigned p = mis_hash(static_cast<unsigned>(v) + 712313887u) ^ flip;
  return (static_cast<unsigned long long>(p) << 32) | static_cast<unsigned>(v);
}

// One lane decides vertex v (short list).  Returns the new state.
IG_DEV int decide_lane(const MisArgs& a, int v, int beg, int end)
{
  // Only neighbours with a larger key can decide v: a smaller one cannot be
  // included while v is undecided (it would need v excluded first), so their
  // states are not read.  Four entries' loads are in flight at a time and
  // evaluated in entry order.
  const unsigned long long kv = mis_key(v, a.flip);
  for (int e = beg; e < end; e += kMisLaneUnroll) {
    int su[kMisLaneUnroll];
#pragma unroll
    for (int q = 0; q < kMisLaneUnroll; q++) {
      const int u = e + q < end ? a.nlist[e + q] : v;
      const bool larger = u != v && mis_key(u, a.flip) > kv;
      su[q] = larger ? ld_relaxed(&a.status[u]) : kExcluded;
    }
#pragma unroll
    for (int q = 0; q < kMisLaneUnroll; q++) {
      if (su[q] == kIncluded) return kExcluded;
      if (su[q] == kUndecided) return kUndecided;
    }
  }
  return kIncluded;
}
// v has just been included: its larger-key neighbours are all excluded
// already, so only the smaller-key ones are read, and the undecided among
// them excluded.  (Storing unconditionally is slower: device-scope stores
// to already excluded hub neighbours cost more than the reads they save.)
IG_DEV void exclude_neighbours_lane(const MisArgs& a, int v, int beg, int end)
{
  const unsigned long long kv = mis_key(v, a.flip);
  for (int e = beg; e < end; e++) {
    const int u = a.nlist[e];
    if (u != v && mis_key(u, a.flip) < kv && ld_relaxed(&a.status[u]) == kUndecided)
      st_relaxed(&a.status[u], kExcluded);
  }
}

// Whole wave decides vertex v (long list), 4 x 64 entries per step; all
// lanes must be converged.
IG_DEV int decide_wave(const MisArgs& a, int v, int beg, int end)
{
  const unsigned long long kv = mis_key(v, a.flip);
  const int lane = lane_id();
  for (int base = beg; base < end; base += 4 * kWave) {
    bool inc = false, hi = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = base + k * kWave + lane;
      if (e < end) {
        const int u = a.nlist[e];
        if (u != v && mis_key(u, a.flip) > kv) {  // smaller keys cannot decide v (see decide_lane)
          const int su = ld_relaxed(&a.status[u]);
          inc |= su == kIncluded;
          hi |= su == kUndecided;
        }
      }
    }
    if (__ballot(inc) != 0) return kExcluded;
    if (__ballot(hi) != 0) return kUndecided;
  }
  return kIncluded;
}

IG_DEV void exclude_neighbours_wave(const MisArgs& a, int v, int beg, int end)
{
  const unsigned long long kv = mis_key(v, a.flip);
  for (int e = beg + lane_id(); e < end; e += kWave) {
    const int u = a.nlist[e];
    if (u != v && mis_key(u, a.flip) < kv && ld_relaxed(&a.status[u]) == kUndecided)  // as exclude_neighbours_lane
      st_relaxed(&a.status[u], kExcluded);
  }
}

__global__ __launch_bounds__(kMisBlock) IG_SGPR_CAP void mis_kernel(MisArgs a)
{
  __shared__ BlockQueue<4096> s_wl;
  __shared__ BlockQueue<1024> s_hwl;
  MisControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int lane = lane_id();
  if (tid == 0) ctl->trace.start = now_ticks();
  for (int i = tid; i < a.nodes; i += stride) st_relaxed(&a.status[i], kUndecided);
  // round 0 long-list worklist: every vertex with >= kMidDeg entries
  {
    const int wpb = blockDim.x / kWave;
    const int nw = (a.nodes + kWave - 1) / kWave;
    bq_reset(s_hwl);
    for (int w = blockIdx.x * wpb + threadIdx.x / kWave; w < nw; w += gridDim.x * wpb) {
      const int v = w * kWave + lane;
      const bool heavy = v < a.nodes && a.nindex[v + 1] - a.nindex[v] >= kMidDeg;
      bq_append(s_hwl, heavy, v, a.hwl[0], &ctl->cnt[0].nm);
    }
    bq_flush(s_hwl, a.hwl[0], &ctl->cnt[0].nm);
  }
  if (!grid_barrier<true>(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    return;
  }
  const int wpb = blockDim.x / kWave;
  const unsigned gwave = blockIdx.x * wpb + threadIdx.x / kWave;
  const unsigned nwaves = gridDim.x * wpb;
  int r = 0;
  for (;; r++) {
    FrontierCounters* ic = &ctl->cnt[r % 3];
    const unsigned n = (r == 0) ? static_cast<unsigned>(a.nodes) : ld_relaxed(&ic->n);
    const unsigned nh = ld_relaxed(&ic->nm);
    if (n == 0 && nh == 0) break;
    trace_level(&ctl->trace, r, n, nh, 0);
    if (tid == 0) counters_clear(&ctl->cnt[(r + 2) % 3]);
    FrontierCounters* oc = &ctl->cnt[(r + 1) % 3];
    const int* in = a.wl[r & 1];
    int* out = a.wl[(r + 1) & 1];
    const int* hin = a.hwl[r & 1];
    int* hout = a.hwl[(r + 1) & 1];
    bq_reset(s_wl);
    bq_reset(s_hwl);
    // short lists: one lane each
    for (unsigned base = blockIdx.x * blockDim.x; base < n; base += stride) {
      const unsigned idx = base + threadIdx.x;
      bool keep = false;
      int v = 0;
      if (idx < n) {
        v = (r == 0) ? static_cast<int>(idx) : ld_relaxed(&in[idx]);
        if (ld_relaxed(&a.status[v]) == kUndecided) {
          const int beg = a.nindex[v], end = a.nindex[v + 1];
          if (end - beg < kMidDeg) {
            const int st = decide_lane(a, v, beg, end);
            if (st == kIncluded) {
              st_relaxed(&a.status[v], kIncluded);
              exclude_neighbours_lane(a, v, beg, end);
            } else if (st == kExcluded) {
              st_relaxed(&a.status[v], kExcluded);
            } else {
              keep = true;
            }
          }
        }
      }
      bq_append(s_wl, keep, v, out, &oc->n);
    }
    // long lists: one wave each
    for (unsigned i = gwave; i < nh; i += nwaves) {
      const int v = ld_relaxed(&hin[i]);
      if (ld_relaxed(&a.status[v]) != kUndecided) continue;
      const int beg = a.nindex[v], end = a.nindex[v + 1];
      const int st = decide_wave(a, v, beg, end);
      if (st == kIncluded) {
        if (lane == 0) st_relaxed(&a.status[v], kIncluded);
        exclude_neighbours_wave(a, v, beg, end);
      } else if (st == kExcluded) {
        if (lane == 0) st_relaxed(&a.status[v], kExcluded);
      }
      bq_append(s_hwl, st == kUndecided && lane == 0, v, hout, &oc->nm);
    }
    bq_flush(s_wl, out, &oc->n);
    bq_flush(s_hwl, hout, &oc->nm);
    if (!grid_barrier<true>(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      return;
    }
  }
  if (tid == 0) ctl->rounds = r;
}

}  // namespace

size_t mis_workspace_bytes(int nodes)
{
  return Carver::round(sizeof(MisControl)) + 4 * Carver::round(static_cast<size_t>(nodes) * 4) + 4096;
}

int mis_run(indigo_ctx* ctx, int nodes, const int* nindex, const int* nlist, int signed_priority, int* status)
{
  Carver cv(ctx->ws);
  MisArgs a{};
  a.ctl = cv.take<MisControl>(1);
  a.wl[0] = cv.take<int>(nodes);
  a.wl[1] = cv.take<int>(nodes);
  a.hwl[0] = cv.take<int>(nodes);
  a.hwl[1] = cv.take<int>(nodes);
  a.nodes = nodes;
  a.flip = signed_priority ? 0x80000000u : 0u;
  a.nindex = nindex;
  a.nlist = nlist;
  a.status = status;
  a.budget = budget_ticks(ctx);
  a.start_ticks = residency_ticks(ctx);
  IG_TRY(zero_async(ctx, a.ctl, sizeof(MisControl)));
  void* args[] = {&a};
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(mis_kernel), kMisBlock, 0, args, 0, nullptr, "mis_kernel", true));
  return INDIGO_OK;
}

int mis_collect(indigo_ctx* ctx, int* rounds)
{
  Carver cv(ctx->ws);
  MisControl* ctl = cv.take<MisControl>(1);
  IG_TRY(ensure_host_pinned(ctx, sizeof(MisControl)));
  MisControl* h = static_cast<MisControl*>(ctx->host_pinned);
  IG_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(MisControl), hipMemcpyDeviceToHost, ctx->stream));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  IG_TRY(check_abort(ctx, reinterpret_cast<const void*>(mis_kernel), h->bar.abort.v, "mis"));
  if (h->status != 0) return set_error(INDIGO_ERR_TIMEOUT, "mis: persistent kernel stopped (status %d)", h->status);
  if (rounds) *rounds = h->rounds;
  store_trace(ctx, h->trace);
  return INDIGO_OK;
}

}  // namespace indigo
