// Breadth-first search (hop distance from one source) on gfx950.
//
// Reference semantics: GPUbfs_vertex / GPUbfs_edge (include/bfs_vertex_cuda.h:61,
// include/bfs_edge_cuda.h:61); dist[v] = number of hops from src, INT_MAX when
// unreachable; the result equals CPUserialDijkstra with unit weights
// (include/bfs_vertex_cuda.h:75-99).
//
// Design: one persistent kernel runs every level; levels are
// separated by the bounded grid barrier (coherent mode: dist, queues and
// counters are only touched with device-scope atomic accesses), so the host
// never round-trips per
// level (the reference copies a goagain flag D2H every level,
// e.g. codes/cuda/bfs-cuda/BFS_CUDA_V_Topo_Push_Determ_IntType_ReadModifyWrite_
// NonPersist_Atomic_Thread_NoBoundsBug2_NoBoundsBug1.cu:98-106).
// Each level is either
//   * top-down: the degree-binned frontier of frontier.hpp (small lists by
//     64-lane waves with the in-wave balancer, mid lists one wave each, huge
//     lists in 2048-entry chunks over all workgroups); a neighbour is claimed
//     with one CAS INT_MAX -> L+1, so each vertex joins the next frontier
//     exactly once; or
//   * bottom-up: every unvisited vertex looks for a parent with dist == L and
//     stops at the first one (Beamer, Asanovic, Patterson, SC'12): short
//     lists one lane each in vertex order, long lists (collected once at
//     start) one wave each with a ballot per 64 entries.
// The direction follows Beamer's heuristic (m_f > m_u / 14 -> bottom-up,
// n_f < V / 24 -> top-down), evaluated identically by every workgroup.
// Small frontiers (high-diameter graphs) are advanced by workgroup 0 alone
// with LDS queues and workgroup barriers only (small_phase).  Vertices without
// neighbours get their level but never enter a frontier.
#include "bfs.hpp"

#include "frontier.hpp"
#include "trace.hpp"

namespace indigo {

namespace {

constexpr int kBfsBlock = 512;
// One-workgroup phase: LDS queue capacity; the phase runs while the frontier's
// degree sum (an upper bound on the next frontier) fits.
constexpr int kSmallCap = 4096;

struct BfsControl {
  GridBarrier bar;
  PaddedU32 small_ack;  // workgroups that have read the counters before a one-workgroup phase
  FrontierCounters cnt[3];
  FrontierCounters heavy;  // static lists of >= kMidDeg entries (bottom-up)
  int levels;
  int bottom_up_levels;
  int small_levels;
  int status;
  int level_base;  // level at which the grid resumes after a one-workgroup phase
  int pad[27];
  LevelTrace trace;
};

struct BfsArgs {
  int nodes;
  int edges;
  int src;
  int allow_bu;
  int allow_small;
  const int* __restrict__ nindex;
  const int* __restrict__ nlist;
  int* dist;
  FrontierQueues fq[2];
  FrontierQueues heavy;
  BfsControl* ctl;
  long long budget;
  long long start_ticks;
};

IG_DEV int degree_of(const int* __restrict__ nindex, int u) { return nindex[u + 1] - nindex[u]; }

// Claims u for level `next` if unvisited; exactly one caller wins.
IG_DEV bool claim(int* dist, int u, int next)
{
  return dist[u] == kInf && cas_relaxed(&dist[u], kInf, next);
}

IG_DEV void top_down(const BfsArgs& a, int L, const FrontierCounters& ic, const FrontierQueues& in,
                     const FrontierQueues& out, FrontierCounters* oc, FrontierEmitter& E)
{
  const int next = L + 1;
  auto visit = [&](int, int e, bool active) {
    bool won = false;
    int u = 0, udeg = 0;
    if (active) {
      u = a.nlist[e];
      won = claim(a.dist, u, next);
      if (won) udeg = degree_of(a.nindex, u);
    }
    frontier_emit_block(E, won && udeg > 0, u, udeg, out, oc);
  };
  expand_huge(in, huge_count(ic.hpack), huge_chunks(ic.hpack), a.nindex, visit);
  expand_mid(in, ic.nm, a.nindex, visit);
  expand_small(in.q, ic.n, a.nindex, visit);
}

IG_DEV void bottom_up(const BfsArgs& a, int L, unsigned heavy_mid, unsigned heavy_huge, const FrontierQueues& out,
                      FrontierCounters* oc, FrontierEmitter& E)
{
  const int next = L + 1;
  const int lane = lane_id();
  // short lists: one lane per vertex, vertex order
  const int stride = gridDim.x * blockDim.x;
  for (int base = blockIdx.x * blockDim.x; base < a.nodes; base += stride) {
    const int v = base + static_cast<int>(threadIdx.x);
    bool found = false;
    int deg = 0;
    if (v < a.nodes && ld_relaxed(&a.dist[v]) == kInf) {
      const int beg = a.nindex[v], end = a.nindex[v + 1];
      deg = end - beg;
      if (deg < kMidDeg) {
        for (int e = beg; e < end; e++) {
          if (ld_relaxed(&a.dist[a.nlist[e]]) == L) {
            found = true;
            break;
          }
        }
        if (found) st_relaxed(&a.dist[v], next);
      }
    }
    frontier_emit_block(E, found, v, deg, out, oc);
  }
  // long lists: one wave per vertex, 64 entries per ballot
  const int wpb = blockDim.x / kWave;
  const unsigned total = heavy_mid + heavy_huge;
  for (unsigned i = blockIdx.x * wpb + threadIdx.x / kWave; i < total; i += gridDim.x * wpb) {
    const int v = i < heavy_mid ? ld_relaxed(&a.heavy.mq[i]) : ld_relaxed(&a.heavy.hq[i - heavy_mid]);
    if (ld_relaxed(&a.dist[v]) != kInf) continue;
    const int beg = a.nindex[v], end = a.nindex[v + 1];
    bool found = false;
    for (int base = beg; base < end; base += kWave) {
      const int e = base + lane;
      if (__ballot(e < end && ld_relaxed(&a.dist[a.nlist[e]]) == L) != 0) {
        found = true;
        break;
      }
    }
    if (found && lane == 0) st_relaxed(&a.dist[v], next);
    frontier_emit_block(E, found && lane == 0, v, end - beg, out, oc);
  }
}

// LDS of the one-workgroup phase (shares the kernel's LDS with the level
// emitter, which is idle meanwhile).
struct SmallLds {
  int q[kSmallCap];      // frontier of the current level, refilled with the next one
  int incl[kSmallCap];   // inclusive prefix of the current frontier's degrees
  int shift[kSmallCap];  // nindex[v] - exclusive prefix: entry k of frontier slot j is nlist[shift[j] + k]
  int wsum[kBfsBlock / kWave];
  int cnt;
  int total;
  int abort_flag;
};

// Workgroup 0 advances levels alone while the frontier's degree sum fits the
// LDS arrays, paying workgroup barriers only.  Entered at level L with the
// frontier in the small global queue (n entries, degree sum <= kSmallCap);
// on leaving (frontier empty, or its degree sum too large) it re-emits the
// frontier of the level it reached into that level's global queues and
// counters, cleared first.  The counters may only be touched once every other
// workgroup has read level L's counters: each of them bumps small_ack
// (release) after deciding on this phase, and workgroup 0 waits for all of
// them (bounded) before the hand-back.  Returns the level reached.
//
// A level is three dependent rounds of global accesses: the frontier's
// index pairs (all at once, one run of slots per thread), then after a
// workgroup prefix sum the list entries spread evenly over all 512 threads,
// then one CAS per entry (INT_MAX -> L + 1, so each vertex is claimed once).
// Claimed vertices join the next frontier whatever their degree; those
// without neighbours contribute no entries and are not handed back.
IG_DEV int small_phase(const BfsArgs& a, int L, unsigned n, BfsControl* ctl, long long deadline, bool* aborted,
                       SmallLds& S)
{
  constexpr int kPer = kSmallCap / kBfsBlock;  // frontier slots per thread
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const int* gin = a.fq[L & 1].q;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) S.q[i] = ld_relaxed(&gin[i]);
  if (threadIdx.x == 0) {
    S.cnt = static_cast<int>(n);
    S.abort_flag = 0;
  }
  __syncthreads();
  *aborted = false;
  int cn = S.cnt;
  for (;;) {
    if (cn == 0) break;
    if ((L & 63) == 0 && threadIdx.x == 0) S.abort_flag = kernel_aborted(&ctl->bar, deadline) ? 1 : 0;
    // round 1: index pairs of this thread's run of frontier slots
    const int i0 = min(cn, static_cast<int>(threadIdx.x) * kPer);
    const int i1 = min(cn, i0 + kPer);
    int beg[kPer], deg[kPer];
    int mine = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const bool ok = i0 + j < i1;
      const int v = ok ? S.q[i0 + j] : 0;
      beg[j] = ok ? a.nindex[v] : 0;
      deg[j] = ok ? a.nindex[v + 1] - beg[j] : 0;
      mine += deg[j];
    }
    const int winc = wave_inclusive_scan(mine);
    if (lane == kWave - 1) S.wsum[wid] = winc;
    __syncthreads();
    if (S.abort_flag) {
      *aborted = true;
      return L;
    }
    int before = 0, total = 0;
    for (int w = 0; w < kBfsBlock / kWave; w++) {
      before += w < wid ? S.wsum[w] : 0;
      total += S.wsum[w];
    }
    if (total > kSmallCap) break;  // S.q still holds level L's frontier
    int run = before + winc - mine;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (i0 + j < i1) {
        S.incl[i0 + j] = run + deg[j];
        S.shift[i0 + j] = beg[j] - run;
        run += deg[j];
      }
    }
    if (threadIdx.x == 0) S.cnt = 0;
    __syncthreads();
    // rounds 2 and 3: list entries, then claims; the next frontier refills S.q
    const int next = L + 1;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int k = j * kBfsBlock + static_cast<int>(threadIdx.x);
      const bool act = k < total;
      int u = 0;
      if (act) {
        int lo = 0, hi = cn - 1;  // owner: the first slot with incl > k
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (S.incl[mid] > k) hi = mid;
          else lo = mid + 1;
        }
        u = a.nlist[S.shift[lo] + k];
      }
      beg[j] = u;  // reuse: the entry's neighbour
      deg[j] = act ? 1 : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) deg[j] = deg[j] && cas_relaxed(&a.dist[beg[j]], kInf, next);
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const unsigned long long m = __ballot(deg[j] != 0);
      if (m != 0) {
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(&S.cnt, __popcll(m));
        b = wave_shfl(b, leader);
        // at most `total` <= kSmallCap claims in all
        if (deg[j]) S.q[b + lanes_below(m)] = beg[j];
      }
    }
    __syncthreads();
    cn = S.cnt;
    L = next;
    __syncthreads();
  }
  // hand the level-L frontier back to the grid, once all others have read
  // the entry level's counters
  if (threadIdx.x == 0) {
    const unsigned others = gridDim.x - 1u;
    while (__hip_atomic_load(&ctl->small_ack.v, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < others) {
      if (kernel_aborted(&ctl->bar, deadline)) {
        S.abort_flag = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(&ctl->small_ack.v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!S.abort_flag)
      for (int k = 0; k < 3; k++) counters_clear(&ctl->cnt[k]);
  }
  __syncthreads();
  if (S.abort_flag) {
    *aborted = true;
    return L;
  }
  for (int base = 0; base < cn; base += blockDim.x) {
    const int i = base + static_cast<int>(threadIdx.x);
    const int v = i < cn ? S.q[i] : 0;
    const int d = i < cn ? degree_of(a.nindex, v) : 0;
    frontier_emit(d > 0, v, d, a.fq[L & 1], &ctl->cnt[L % 3]);
  }
  return L;
}

__global__ __launch_bounds__(kBfsBlock) void bfs_kernel(BfsArgs a)
{
  // the set-up's heavy-list collection needs a larger mid buffer than the
  // level emitter (no spills to the global counter); they share the LDS
  __shared__ union {
    FrontierEmitter emit;
    HeavyCollector collect;
    SmallLds small;
  } s_lds;
  FrontierEmitter& s_emit = s_lds.emit;
  BfsControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  if (tid == 0) ctl->trace.start = now_ticks();
  for (int i = tid; i < a.nodes; i += stride) st_relaxed(&a.dist[i], (i == a.src) ? 0 : kInf);
  if (a.allow_bu) collect_heavy(s_lds.collect, a.nodes, a.nindex, a.heavy, &ctl->heavy);
  if (blockIdx.x == 0 && threadIdx.x < kWave) {
    const int deg = degree_of(a.nindex, a.src);
    frontier_emit(threadIdx.x == 0 && deg > 0, a.src, deg, a.fq[0], &ctl->cnt[0]);
  }
  if (!grid_barrier<true>(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    return;
  }
  const unsigned heavy_mid = ld_relaxed(&ctl->heavy.nm);
  const unsigned heavy_huge = huge_count(ld_relaxed(&ctl->heavy.hpack));
  bool bu = false;
  unsigned long long seen_edges = 0;
  int bu_levels = 0;
  int L = 0;
  for (;;) {
    const FrontierCounters ic = counters_load(&ctl->cnt[L % 3]);
    const unsigned long long nf = static_cast<unsigned long long>(ic.n) + ic.nm + huge_count(ic.hpack);
    if (nf == 0) break;
    seen_edges += ic.m;
    if (a.allow_bu) {
      const unsigned long long e = static_cast<unsigned long long>(a.edges);
      const unsigned long long mu = e > seen_edges ? e - seen_edges : 0ull;
      if (!bu) {
        if (ic.m * 14ull > mu) bu = true;
      } else if (nf * 24ull < static_cast<unsigned long long>(a.nodes)) {
        bu = false;
      }
    }
    const bool small = !bu && a.allow_small && ic.nm == 0 && ic.hpack == 0 &&
                       ic.m <= static_cast<unsigned long long>(kSmallCap);
    trace_level(&ctl->trace, L, static_cast<unsigned>(nf), ic.m, small ? 2 : (bu ? 1 : 0));
    if (small) {
      // one-workgroup phase; the other workgroups acknowledge that they have
      // read this level's counters and wait at the next barrier
      if (blockIdx.x != 0 && threadIdx.x == 0)
        __hip_atomic_fetch_add(&ctl->small_ack.v, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (blockIdx.x == 0) {
        bool aborted = false;
        const int L2 = small_phase(a, L, ic.n, ctl, deadline, &aborted, s_lds.small);
        if (threadIdx.x == 0) {
          st_relaxed(&ctl->level_base, L2);
          ctl->small_levels += L2 - L;
        }
      }
      if (!grid_barrier<true>(&ctl->bar, deadline)) {
        if (tid == 0) ctl->status = 1;
        return;
      }
      const int L2 = ld_relaxed(&ctl->level_base);
      if (L2 == L) {  // cannot happen: the phase always advances at least one level
        if (tid == 0) ctl->status = 2;
        return;
      }
      L = L2;
      continue;
    }
    if (tid == 0) counters_clear(&ctl->cnt[(L + 2) % 3]);
    const FrontierQueues& in = a.fq[L & 1];
    const FrontierQueues& out = a.fq[(L + 1) & 1];
    FrontierCounters* oc = &ctl->cnt[(L + 1) % 3];
    fe_reset(s_emit);
    if (bu) {
      bottom_up(a, L, heavy_mid, heavy_huge, out, oc, s_emit);
      bu_levels++;
    } else {
      top_down(a, L, ic, in, out, oc, s_emit);
    }
    fe_flush(s_emit, out, oc);
    if (!grid_barrier<true>(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      return;
    }
    L++;
  }
  if (tid == 0) {
    ctl->levels = L;
    ctl->bottom_up_levels = bu_levels;
  }
}

}  // namespace

size_t bfs_workspace_bytes(int nodes)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(BfsControl)) + 12 * Carver::round(v * 4) + 4096;
}

int bfs_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int src, int* dist,
            BfsStats* stats)
{
  IG_TRY(ensure_workspace(ctx, bfs_workspace_bytes(nodes)));
  IG_TRY(ensure_host_pinned(ctx, sizeof(BfsControl)));
  Carver cv(ctx->ws);
  BfsArgs a{};
  a.ctl = cv.take<BfsControl>(1);
  for (int k = 0; k < 3; k++) {
    FrontierQueues& f = k < 2 ? a.fq[k] : a.heavy;
    f.q = cv.take<int>(nodes);
    f.mq = cv.take<int>(nodes);
    f.hq = cv.take<int>(nodes);
    f.hoff = cv.take<unsigned>(nodes);
  }
  a.nodes = nodes;
  a.edges = edges;
  a.src = src;
  a.allow_bu = ctx->hints[INDIGO_HINT_BFS_DIRECTION] != 1;
  a.allow_small = ctx->hints[INDIGO_HINT_SMALL_FRONTIER] != 1;
  a.nindex = nindex;
  a.nlist = nlist;
  a.dist = dist;
  a.budget = budget_ticks(ctx);
  a.start_ticks = residency_ticks(ctx);
  IG_TRY(zero_async(ctx, a.ctl, sizeof(BfsControl)));
  void* args[] = {&a};
  int grid = 0;
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(bfs_kernel), kBfsBlock, 0, args, 0, &grid, "bfs_kernel"));
  if (stats) {
    stats->launches = 2;
    stats->grid = grid;
  }
  return INDIGO_OK;
}

int bfs_collect(indigo_ctx* ctx, BfsStats* stats)
{
  Carver cv(ctx->ws);
  BfsControl* ctl = cv.take<BfsControl>(1);
  BfsControl* h = static_cast<BfsControl*>(ctx->host_pinned);
  IG_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(BfsControl), hipMemcpyDeviceToHost, ctx->stream));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  IG_TRY(check_abort(ctx, reinterpret_cast<const void*>(bfs_kernel), h->bar.abort.v, "bfs"));
  if (h->status != 0)
    return set_error(INDIGO_ERR_TIMEOUT, "bfs: persistent kernel stopped (status %d) before finishing", h->status);
  if (stats) {
    stats->levels = h->levels;
    stats->bottom_up_levels = h->bottom_up_levels;
    stats->small_levels = h->small_levels;
  }
  store_trace(ctx, h->trace);
  return INDIGO_OK;
}

}  // namespace indigo
