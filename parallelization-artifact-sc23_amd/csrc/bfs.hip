// Breadth-first search (hop distance from one source) on gfx950.
//
// Reference semantics: GPUbfs_vertex / GPUbfs_edge (include/bfs_vertex_cuda.h:61,
// include/bfs_edge_cuda.h:61); dist[v] = number of hops from src, INT_MAX when
// unreachable; the result equals CPUserialDijkstra with unit weights
// (include/bfs_vertex_cuda.h:75-99).
//
// Design: one cooperative persistent kernel runs every level; levels are
// separated by the bounded grid barrier, so the host never round-trips per
// level (the reference copies a goagain flag D2H every level,
// e.g. codes/cuda/bfs-cuda/BFS_CUDA_V_Topo_Push_Determ_IntType_ReadModifyWrite_
// NonPersist_Atomic_Thread_NoBoundsBug2_NoBoundsBug1.cu:98-106).
// Each level is either
//   * top-down: the degree-binned frontier (normal queue by 64-lane waves with
//     an in-wave edge balancer, huge lists split in 2048-edge chunks over all
//     workgroups); a neighbour is claimed with one CAS INT_MAX -> L+1 so each
//     vertex enters the next frontier exactly once; or
//   * bottom-up: every unvisited vertex scans its list for a parent with
//     dist == L and stops at the first one (Beamer's direction-optimising BFS,
//     SC'12), which reads far fewer than E entries on low-diameter graphs.
// The direction switches on Beamer's heuristic (m_f > m_u / 14 -> bottom-up,
// n_f < V / 24 -> top-down), evaluated identically by every workgroup.
// Small frontiers (high-diameter graphs) are advanced by workgroup 0 alone
// with LDS-resident queues and workgroup barriers only; see bfs_small_phase.
#include "bfs.hpp"

#include "frontier.hpp"

namespace indigo {

namespace {

constexpr int kBfsBlock = 512;
// One-workgroup phase: frontier capacity held in LDS and the frontier size
// below which the grid hands a level over to workgroup 0.
constexpr int kSmallCap = 4096;
constexpr int kSmallEnter = 2048;

struct BfsControl {
  GridBarrier bar;
  FrontierCounters cnt[3];
  int levels;
  int bottom_up_levels;
  int small_levels;
  int status;
  int level_base;  // level at which the grid resumes after a one-workgroup phase
  int pad[27];
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

IG_DEV void top_down(const BfsArgs& a, int L, unsigned n, unsigned nh, unsigned nch, const FrontierQueues& in,
                     const FrontierQueues& out, FrontierCounters* oc)
{
  __shared__ int s_owner;
  const int lane = lane_id();
  const int next = L + 1;
  // huge adjacency lists: one workgroup per kHugeChunk-edge item
  for (unsigned item = blockIdx.x; item < nch; item += gridDim.x) {
    if (threadIdx.x == 0) s_owner = huge_owner(in.hoff, static_cast<int>(nh), item);
    __syncthreads();
    const int j = s_owner;
    __syncthreads();
    const int v = in.hq[j];
    const int beg = a.nindex[v] + static_cast<int>(item - in.hoff[j]) * kHugeChunk;
    const int end = min(a.nindex[v + 1], beg + kHugeChunk);
    for (int base = beg; base < end; base += blockDim.x) {
      const int e = base + static_cast<int>(threadIdx.x);
      bool won = false;
      int u = 0, udeg = 0;
      if (e < end) {
        u = a.nlist[e];
        won = claim(a.dist, u, next);
        if (won) udeg = degree_of(a.nindex, u);
      }
      frontier_emit(won, u, udeg, out, oc);
    }
  }
  // normal queue: one wave per 64 frontier vertices, edges balanced in-wave
  const int wpb = blockDim.x / kWave;
  const unsigned gwave = blockIdx.x * wpb + threadIdx.x / kWave;
  const unsigned nwaves = gridDim.x * wpb;
  for (unsigned c = gwave; c * kWave < n; c += nwaves) {
    const unsigned idx = c * kWave + lane;
    const bool valid = idx < n;
    const int v = valid ? in.q[idx] : 0;
    const int beg = valid ? a.nindex[v] : 0;
    const int deg = valid ? a.nindex[v + 1] - beg : 0;
    const int incl = wave_inclusive_scan(deg);
    const int total = wave_shfl(incl, kWave - 1);
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int obeg = wave_shfl(beg, o);
      const int oexcl = wave_shfl(incl - deg, o);
      bool won = false;
      int u = 0, udeg = 0;
      if (k < total) {
        u = a.nlist[obeg + (k - oexcl)];
        won = claim(a.dist, u, next);
        if (won) udeg = degree_of(a.nindex, u);
      }
      frontier_emit(won, u, udeg, out, oc);
    }
  }
}

IG_DEV void bottom_up(const BfsArgs& a, int L, const FrontierQueues& out, FrontierCounters* oc)
{
  const int next = L + 1;
  const int stride = gridDim.x * blockDim.x;
  for (int base = blockIdx.x * blockDim.x; base < a.nodes; base += stride) {
    const int v = base + static_cast<int>(threadIdx.x);
    bool found = false;
    int deg = 0;
    if (v < a.nodes && a.dist[v] == kInf) {
      const int beg = a.nindex[v], end = a.nindex[v + 1];
      deg = end - beg;
      for (int e = beg; e < end; e++) {
        if (a.dist[a.nlist[e]] == L) {
          found = true;
          break;
        }
      }
      if (found) a.dist[v] = next;
    }
    frontier_emit(found, v, deg, out, oc);
  }
}

// Workgroup 0 advances levels alone while the frontier fits in LDS: each
// level costs workgroup barriers only.  Entered with the level-L frontier in
// the normal global queue `in` (n entries, no huge entries); leaves when the
// frontier is empty or outgrows kSmallCap, writing it back to the global
// queue of the level it stopped at.  Returns the level reached.
IG_DEV int small_phase(const BfsArgs& a, int L, unsigned n, const FrontierQueues* fq, GridBarrier* bar,
                       long long deadline, unsigned* out_n, bool* aborted)
{
  __shared__ int s_q[2][kSmallCap];
  __shared__ int s_cnt[2];
  __shared__ int s_overflow;
  const int lane = lane_id();
  const int* gin = fq[L & 1].q;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) s_q[0][i] = gin[i];
  if (threadIdx.x == 0) {
    s_cnt[0] = static_cast<int>(n);
    s_cnt[1] = 0;
    s_overflow = 0;
  }
  __syncthreads();
  int cur = 0;
  *aborted = false;
  for (;;) {
    const int cn = s_cnt[cur];
    if (cn == 0 || cn > kSmallEnter) break;
    if ((L & 63) == 0 && threadIdx.x == 0) s_overflow = kernel_aborted(bar, deadline) ? 2 : 0;
    __syncthreads();
    if (s_overflow == 2) {
      *aborted = true;
      return L;
    }
    const int nxt = cur ^ 1;
    const int next = L + 1;
    // expand: one wave per 64 frontier vertices with the in-wave balancer
    const int wpb = blockDim.x / kWave;
    for (int c = threadIdx.x / kWave; c * kWave < cn; c += wpb) {
      const int idx = c * kWave + lane;
      const bool valid = idx < cn;
      const int v = valid ? s_q[cur][idx] : 0;
      const int beg = valid ? a.nindex[v] : 0;
      const int deg = valid ? a.nindex[v + 1] - beg : 0;
      const int incl = wave_inclusive_scan(deg);
      const int total = wave_shfl(incl, kWave - 1);
      for (int base = 0; base < total; base += kWave) {
        const int k = base + lane;
        const int o = min(wave_owner(incl, k), kWave - 1);
        const int obeg = wave_shfl(beg, o);
        const int oexcl = wave_shfl(incl - deg, o);
        bool won = false;
        int u = 0;
        if (k < total) {
          u = a.nlist[obeg + (k - oexcl)];
          won = claim(a.dist, u, next);
        }
        const unsigned long long m = __ballot(won);
        if (m != 0) {
          const int leader = __ffsll(static_cast<long long>(m)) - 1;
          int b = 0;
          if (lane == leader) b = atomicAdd(&s_cnt[nxt], __popcll(m));
          b = wave_shfl(b, leader);
          const int slot = b + lanes_below(m);
          if (won) {
            if (slot < kSmallCap) s_q[nxt][slot] = u;
            else fq[next & 1].q[slot] = u;  // spill: written straight to the global queue
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_cnt[cur] = 0;
    cur = nxt;
    L = next;
    __syncthreads();
  }
  // write the frontier of level L back to the global queue for that level
  const int cn = s_cnt[cur];
  int* gout = fq[L & 1].q;
  for (int i = threadIdx.x; i < cn && i < kSmallCap; i += blockDim.x) gout[i] = s_q[cur][i];
  *out_n = static_cast<unsigned>(cn);
  return L;
}

__global__ __launch_bounds__(kBfsBlock) void bfs_kernel(BfsArgs a)
{
  BfsControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  for (int i = tid; i < a.nodes; i += stride) a.dist[i] = (i == a.src) ? 0 : kInf;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int deg = degree_of(a.nindex, a.src);
    FrontierCounters* c0 = &ctl->cnt[0];
    if (deg >= kHugeDeg) {
      a.fq[0].hq[0] = a.src;
      a.fq[0].hoff[0] = 0;
      c0->hpack = (static_cast<unsigned long long>((deg + kHugeChunk - 1) / kHugeChunk) << 32) | 1ull;
    } else {
      a.fq[0].q[0] = a.src;
      c0->n = 1;
    }
    c0->m = static_cast<unsigned long long>(deg);
    ctl->level_base = 0;
  }
  if (!grid_barrier(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    return;
  }
  bool bu = false;
  unsigned long long seen_edges = 0;
  int bu_levels = 0, small_levels = 0;
  int L = 0;
  for (;;) {
    FrontierCounters* ic = &ctl->cnt[L % 3];
    const unsigned n = ld_relaxed(&ic->n);
    const unsigned long long hp = ld_relaxed(&ic->hpack);
    const unsigned long long mf = ld_relaxed(&ic->m);
    const unsigned nh = huge_count(hp), nch = huge_chunks(hp);
    const unsigned long long nf = static_cast<unsigned long long>(n) + nh;
    if (nf == 0) break;
    seen_edges += mf;
    if (a.allow_bu) {
      const unsigned long long e = static_cast<unsigned long long>(a.edges);
      const unsigned long long mu = e > seen_edges ? e - seen_edges : 0ull;
      if (!bu) {
        if (mf * 14ull > mu) bu = true;
      } else if (nf * 24ull < static_cast<unsigned long long>(a.nodes)) {
        bu = false;
      }
    }
    if (!bu && a.allow_small && nh == 0 && n <= static_cast<unsigned>(kSmallEnter)) {
      // one-workgroup phase; the other workgroups wait at the next barrier
      if (blockIdx.x == 0) {
        unsigned out_n = 0;
        bool aborted = false;
        const int L2 = small_phase(a, L, n, a.fq, &ctl->bar, deadline, &out_n, &aborted);
        if (threadIdx.x == 0) {
          if (!aborted) {
            FrontierCounters* oc = &ctl->cnt[L2 % 3];
            FrontierCounters* z1 = &ctl->cnt[(L2 + 1) % 3];
            FrontierCounters* z2 = &ctl->cnt[(L2 + 2) % 3];
            z1->n = 0; z1->hpack = 0; z1->m = 0;
            z2->n = 0; z2->hpack = 0; z2->m = 0;
            oc->n = out_n;
            oc->hpack = 0;
            // degree sum of the hand-back frontier is not tracked; a zero keeps
            // the direction heuristic in top-down mode for the next level
            oc->m = 0;
          }
          ctl->level_base = L2;
          ctl->small_levels += L2 - L;
        }
      }
      if (!grid_barrier(&ctl->bar, deadline)) {
        if (tid == 0) ctl->status = 1;
        return;
      }
      const int L2 = ld_relaxed(&ctl->level_base);
      small_levels += L2 - L;
      if (L2 == L) {
        // no progress possible in the small phase (should not happen)
        if (tid == 0) ctl->status = 2;
        return;
      }
      L = L2;
      continue;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      FrontierCounters* z = &ctl->cnt[(L + 2) % 3];
      z->n = 0;
      z->hpack = 0;
      z->m = 0;
    }
    const FrontierQueues& in = a.fq[L & 1];
    const FrontierQueues& out = a.fq[(L + 1) & 1];
    FrontierCounters* oc = &ctl->cnt[(L + 1) % 3];
    if (bu) {
      bottom_up(a, L, out, oc);
      bu_levels++;
    } else {
      top_down(a, L, n, nh, nch, in, out, oc);
    }
    if (!grid_barrier(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      return;
    }
    L++;
  }
  if (tid == 0) {
    ctl->levels = L;
    ctl->bottom_up_levels = bu_levels;
  }
  (void)small_levels;
}

}  // namespace

size_t bfs_workspace_bytes(int nodes)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(BfsControl)) + 6 * Carver::round(v * 4) + 4096;
}

int bfs_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int src, int* dist,
            BfsStats* stats)
{
  IG_TRY(ensure_workspace(ctx, bfs_workspace_bytes(nodes)));
  IG_TRY(ensure_host_pinned(ctx, sizeof(BfsControl)));
  Carver cv(ctx->ws);
  BfsArgs a{};
  a.ctl = cv.take<BfsControl>(1);
  for (int k = 0; k < 2; k++) {
    a.fq[k].q = cv.take<int>(nodes);
    a.fq[k].hq = cv.take<int>(nodes);
    a.fq[k].hoff = cv.take<unsigned>(nodes);
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
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(bfs_kernel), kBfsBlock, 0, args, 0, &grid));
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
  return INDIGO_OK;
}

}  // namespace indigo
