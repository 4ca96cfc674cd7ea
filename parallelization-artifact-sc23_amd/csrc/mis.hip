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
// Design: one cooperative persistent kernel; round r processes a worklist of
// still-undecided vertices (all vertices in round 0) and appends the ones that
// stay undecided to the next worklist (wave-aggregated).  Decisions are made
// in place: a decided state (1 included, 2 excluded) never changes, so a
// vertex that reads a neighbour's final state may act on it in the same round
// and a stale "undecided" read only delays it.  A vertex is included only when
// every neighbour with a larger key is excluded, and excluded as soon as a
// neighbour is included: exactly the greedy order's fixed point.  Lists of
// fewer than 64 entries are scanned by one lane; longer ones by the whole
// wave, 64 entries per step with ballots.  Keys are (priority ^ sign flip,
// id) compared as one 64-bit unsigned value; the priority is recomputed from
// the hash instead of being stored.  Self loops are ignored.
#include "mis.hpp"

#include "device_common.hpp"

namespace indigo {

namespace {

constexpr int kMisBlock = 512;
constexpr int kWaveScan = 64;  // lists at least this long are scanned by a whole wave

constexpr int kUndecided = 0, kIncluded = 1, kExcluded = 2;

struct MisControl {
  GridBarrier bar;
  PaddedU32 cnt[3];
  int rounds;
  int status;
  int pad[30];
};

struct MisArgs {
  int nodes;
  unsigned flip;
  const int* __restrict__ nindex;
  const int* __restrict__ nlist;
  int* status;
  int* wl[2];
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
  const unsigned long long kv = mis_key(v, a.flip);
  for (int e = beg; e < end; e++) {
    const int u = a.nlist[e];
    if (u == v) continue;
    const int su = ld_relaxed(&a.status[u]);
    if (su == kIncluded) return kExcluded;
    if (su == kUndecided && mis_key(u, a.flip) > kv) return kUndecided;
  }
  return kIncluded;
}

IG_DEV void exclude_neighbours_lane(const MisArgs& a, int v, int beg, int end)
{
  for (int e = beg; e < end; e++) {
    const int u = a.nlist[e];
    if (u != v && ld_relaxed(&a.status[u]) == kUndecided) st_relaxed(&a.status[u], kExcluded);
  }
}

// Whole wave decides vertex v (long list); all lanes must be converged.
IG_DEV int decide_wave(const MisArgs& a, int v, int beg, int end)
{
  const unsigned long long kv = mis_key(v, a.flip);
  const int lane = lane_id();
  for (int base = beg; base < end; base += kWave) {
    const int e = base + lane;
    bool inc = false, hi = false;
    if (e < end) {
      const int u = a.nlist[e];
      if (u != v) {
        const int su = ld_relaxed(&a.status[u]);
        inc = su == kIncluded;
        hi = su == kUndecided && mis_key(u, a.flip) > kv;
      }
    }
    if (__ballot(inc) != 0) return kExcluded;
    if (__ballot(hi) != 0) return kUndecided;
  }
  return kIncluded;
}

IG_DEV void exclude_neighbours_wave(const MisArgs& a, int v, int beg, int end)
{
  for (int e = beg + lane_id(); e < end; e += kWave) {
    const int u = a.nlist[e];
    if (u != v && ld_relaxed(&a.status[u]) == kUndecided) st_relaxed(&a.status[u], kExcluded);
  }
}

__global__ __launch_bounds__(kMisBlock) void mis_kernel(MisArgs a)
{
  MisControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  for (int i = tid; i < a.nodes; i += gridDim.x * blockDim.x) a.status[i] = kUndecided;
  if (!grid_barrier(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    return;
  }
  const int wpb = blockDim.x / kWave;
  const unsigned gwave = blockIdx.x * wpb + threadIdx.x / kWave;
  const unsigned nwaves = gridDim.x * wpb;
  int r = 0;
  for (;; r++) {
    const unsigned n = (r == 0) ? static_cast<unsigned>(a.nodes) : ld_relaxed(&ctl->cnt[r % 3].v);
    if (n == 0) break;
    if (tid == 0) st_relaxed(&ctl->cnt[(r + 2) % 3].v, 0u);
    const int* in = a.wl[r & 1];
    int* out = a.wl[(r + 1) & 1];
    unsigned* oc = &ctl->cnt[(r + 1) % 3].v;
    for (unsigned c = gwave; c * kWave < n; c += nwaves) {
      const unsigned idx = c * kWave + lane;
      const bool valid = idx < n;
      const int v = valid ? (r == 0 ? static_cast<int>(idx) : in[idx]) : 0;
      int state = kExcluded;  // "nothing to do" for invalid lanes
      int beg = 0, end = 0;
      if (valid) {
        state = ld_relaxed(&a.status[v]);
        if (state == kUndecided) {
          beg = a.nindex[v];
          end = a.nindex[v + 1];
          if (end - beg < kWaveScan) {
            state = decide_lane(a, v, beg, end);
            if (state == kIncluded) {
              st_relaxed(&a.status[v], kIncluded);
              exclude_neighbours_lane(a, v, beg, end);
            } else if (state == kExcluded) {
              st_relaxed(&a.status[v], kExcluded);
            }
          }
        }
      }
      // long lists: the whole wave handles them one vertex at a time
      unsigned long long big = __ballot(valid && state == kUndecided && end - beg >= kWaveScan);
      while (big != 0) {
        const int l = __ffsll(static_cast<long long>(big)) - 1;
        big &= big - 1;
        const int bv = wave_shfl(v, l);
        const int bb = wave_shfl(beg, l);
        const int be = wave_shfl(end, l);
        const int s = decide_wave(a, bv, bb, be);
        if (s == kIncluded) {
          if (lane == 0) st_relaxed(&a.status[bv], kIncluded);
          exclude_neighbours_wave(a, bv, bb, be);
        } else if (s == kExcluded) {
          if (lane == 0) st_relaxed(&a.status[bv], kExcluded);
        }
        if (lane == l) state = s;
      }
      wave_append(valid && state == kUndecided, v, out, oc);
    }
    if (!grid_barrier(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      return;
    }
  }
  if (tid == 0) ctl->rounds = r;
}

}  // namespace

size_t mis_workspace_bytes(int nodes)
{
  return Carver::round(sizeof(MisControl)) + 2 * Carver::round(static_cast<size_t>(nodes) * 4) + 4096;
}

int mis_run(indigo_ctx* ctx, int nodes, const int* nindex, const int* nlist, int signed_priority, int* status)
{
  Carver cv(ctx->ws);
  MisArgs a{};
  a.ctl = cv.take<MisControl>(1);
  a.wl[0] = cv.take<int>(nodes);
  a.wl[1] = cv.take<int>(nodes);
  a.nodes = nodes;
  a.flip = signed_priority ? 0x80000000u : 0u;
  a.nindex = nindex;
  a.nlist = nlist;
  a.status = status;
  a.budget = budget_ticks(ctx);
  a.start_ticks = residency_ticks(ctx);
  IG_TRY(zero_async(ctx, a.ctl, sizeof(MisControl)));
  void* args[] = {&a};
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(mis_kernel), kMisBlock, 0, args, 0, nullptr));
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
  return INDIGO_OK;
}

}  // namespace indigo
