// Single-source shortest paths (int32 weights and distances) on gfx950.
//
// Reference semantics: GPUsssp_vertex / GPUsssp_edge (include/sssp_vertex_cuda.h:60,
// include/sssp_edge_cuda.h:60): dist[src] = 0, INT_MAX when unreachable,
// relaxations d = dist[v] + w in int arithmetic; results equal the serial
// Dijkstra of sssp_vertex_cuda.h:74-99.
//
// Design: one persistent kernel (coherent-mode barriers: dist,
// stamps, queues, piles and counters only see device-scope atomic accesses)
// running near-far
// label-correcting SSSP (Davidson et al., IPDPS'14):
//   * the near frontier holds vertices whose distance improved below the
//     split T; it is expanded with the degree-binned machinery of
//     frontier.hpp (64-lane in-wave edge balancing, huge lists split over
//     workgroups); each edge relaxes with atomicMin and, on improvement, the
//     target joins the next near frontier (deduplicated per iteration with an
//     exchange on stamp[u]) or, if its distance is >= T, the far pile
//     (deduplicated per epoch with farstamp[u]) while far_min tracks the
//     smallest parked distance;
//   * far piles rotate over three buffers: epoch e files into pile e % 3 and
//     re-files into pile (e + 1) % 3 while pile (e + 2) % 3, consumed one
//     barrier earlier, is cleared - no buffer ever has two roles at once;
//   * when the near frontier empties, the split moves to
//     max(T, far_min) + delta and one pass re-files the far pile: entries
//     below the new split join the near frontier, settled entries (distance
//     below the old split, already expanded at that distance) are dropped,
//     the rest stay parked;
//   * vertices are expanded with their current distance (re-read at
//     expansion), so stale duplicates only cost work, never correctness.
// delta defaults to a sampled mean weight x 32 / mean degree (at least 1).
#include "sssp.hpp"

#include "frontier.hpp"
#include "trace.hpp"

namespace indigo {

namespace {

constexpr int kSsspBlock = 512;

struct SsspControl {
  GridBarrier bar;
  FrontierCounters cnt[3];
  PaddedU32 far_cnt[3];
  PaddedU32 far_min[3];
  int iterations;
  int epochs;
  int status;
  int delta;
  int pad[28];
  LevelTrace trace;
};

struct SsspArgs {
  int nodes;
  int edges;
  int src;
  int delta_hint;
  const int* __restrict__ nindex;
  const int* __restrict__ nlist;
  const int* __restrict__ eweight;
  int* dist;
  int* stamp;
  int* farstamp;
  int* far[3];
  FrontierQueues fq[2];
  SsspControl* ctl;
  long long budget;
  long long start_ticks;
};

IG_DEV int degree_of(const int* __restrict__ nindex, int u) { return nindex[u + 1] - nindex[u]; }

// Relaxes one CSR entry e of v (distance dv) when `active`: an improvement
// below the split T joins the next near frontier (deduplicated per iteration
// by stamp), one at or above T is parked in the far pile (deduplicated per
// epoch by farstamp).  Wave-collective: all lanes converged.
// Workgroup-aggregated emission state of one SSSP phase.
struct SsspEmit {
  FrontierEmitter near;
  BlockQueue<2048> far;
  unsigned far_min;
};

IG_DEV void se_reset(SsspEmit& S)
{
  if (threadIdx.x == 0) S.far_min = static_cast<unsigned>(kInf);
  fe_reset(S.near);
  bq_reset(S.far);
}

IG_DEV void se_park(SsspEmit& S, bool park, int u, int d, int* far, unsigned* far_cnt)
{
  bq_append(S.far, park, u, far, far_cnt);
  const int m = wave_min(park ? d : kInf);
  if (lane_id() == 0 && m != kInf) atomicMin(&S.far_min, static_cast<unsigned>(m));
}

IG_DEV void se_flush(SsspEmit& S, const FrontierQueues& out, FrontierCounters* oc, int* far, unsigned* far_cnt,
                     unsigned* far_min)
{
  fe_flush(S.near, out, oc);
  bq_flush(S.far, far, far_cnt);
  if (threadIdx.x == 0 && S.far_min != static_cast<unsigned>(kInf)) atomic_min_relaxed(far_min, S.far_min);
  __syncthreads();
}

IG_DEV void relax(const SsspArgs& a, int dv, int e, bool active, int T, int it_stamp, int epoch, int* far,
                  unsigned* far_cnt, SsspEmit& S, const FrontierQueues& out, FrontierCounters* oc)
{
  bool near = false, park = false;
  int u = 0, nd = 0, udeg = 0;
  if (active) {
    u = a.nlist[e];
    nd = dv + a.eweight[e];
    // device-scope read: a plain load may see a stale L2 copy of a distance
    // another XCD has lowered, and let a useless atomic through
    if (nd < ld_relaxed(&a.dist[u])) {
      const int old = atomic_min_relaxed(&a.dist[u], nd);
      if (nd < old) {
        udeg = degree_of(a.nindex, u);
        if (udeg > 0) {
          if (nd < T)
            near = __hip_atomic_exchange(&a.stamp[u], it_stamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != it_stamp;
          else
            park = __hip_atomic_exchange(&a.farstamp[u], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch;
        }
      }
    }
  }
  frontier_emit_block(S.near, near, u, udeg, out, oc);
  se_park(S, park, u, nd, far, far_cnt);
}

__global__ __launch_bounds__(kSsspBlock) void sssp_kernel(SsspArgs a)
{
  __shared__ SsspEmit s_emit;
  SsspControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int lane = lane_id();
  if (tid == 0) ctl->trace.start = now_ticks();
  for (int i = tid; i < a.nodes; i += stride) {
    st_relaxed(&a.dist[i], (i == a.src) ? 0 : kInf);
    st_relaxed(&a.stamp[i], 0);
    st_relaxed(&a.farstamp[i], 0);
  }
  if (blockIdx.x == 0) {
    // delta: sampled mean weight x 32 / mean degree
    __shared__ long long s_wsum[kSsspBlock / kWave];
    __shared__ int s_ns[kSsspBlock / kWave];
    long long wsum = 0;
    int ns = 0;
    if (a.delta_hint <= 0 && a.edges > 0) {
      const int samples = a.edges < 4096 ? a.edges : 4096;
      for (int k = threadIdx.x; k < samples; k += blockDim.x) {
        const long long e = static_cast<long long>(k) * a.edges / samples;
        wsum += a.eweight[e];
        ns++;
      }
    }
    wsum = wave_sum(wsum);
    ns = wave_sum(ns);
    if (lane == 0) {
      s_wsum[threadIdx.x / kWave] = wsum;
      s_ns[threadIdx.x / kWave] = ns;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int delta = a.delta_hint;
      if (delta <= 0) {
        long long ws = 0;
        int cnt = 0;
        for (int w = 0; w < kSsspBlock / kWave; w++) {
          ws += s_wsum[w];
          cnt += s_ns[w];
        }
        const double mean_w = cnt > 0 ? static_cast<double>(ws) / cnt : 1.0;
        const double mean_d = static_cast<double>(a.edges) / a.nodes;
        double d = mean_w * 32.0 / (mean_d > 1.0 ? mean_d : 1.0);
        delta = d < 1.0 ? 1 : (d > 1e9 ? 1000000000 : static_cast<int>(d));
      }
      st_relaxed(&ctl->delta, delta);
      for (int k = 0; k < 3; k++) st_relaxed(&ctl->far_min[k].v, static_cast<unsigned>(kInf));
    }
    __syncthreads();
    if (threadIdx.x < kWave) {
      const int deg = degree_of(a.nindex, a.src);
      frontier_emit(threadIdx.x == 0 && deg > 0, a.src, deg, a.fq[0], &ctl->cnt[0]);
    }
  }
  if (!grid_barrier<true>(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    return;
  }
  const int delta = ld_relaxed(&ctl->delta);
  int T = delta;  // split: near holds distances < T
  int epoch = 1;  // far pile generation (farstamp value); far[epoch % 3] is current
  int it = 0;     // near iterations (stamp values start at 1)
  int L = 0;      // counter rotation index
  for (;;) {
    FrontierCounters* ic = &ctl->cnt[L % 3];
    const unsigned n = ld_relaxed(&ic->n);
    const unsigned nm = ld_relaxed(&ic->nm);
    const unsigned long long hp = ld_relaxed(&ic->hpack);
    const unsigned nh = huge_count(hp), nch = huge_chunks(hp);
    FrontierCounters* oc = &ctl->cnt[(L + 1) % 3];
    const FrontierQueues& in = a.fq[L & 1];
    const FrontierQueues& out = a.fq[(L + 1) & 1];
    if (tid == 0) counters_clear(&ctl->cnt[(L + 2) % 3]);
    int* far = a.far[epoch % 3];
    unsigned* far_cnt = &ctl->far_cnt[epoch % 3].v;
    unsigned* far_min = &ctl->far_min[epoch % 3].v;
    const bool split = static_cast<unsigned long long>(n) + nm + nh == 0;
    trace_level(&ctl->trace, L, split ? ld_relaxed(far_cnt) : n + nm + nh, split ? 0ull : ld_relaxed(&ic->m),
                split ? 3 : 0);
    if (split) {
      // near frontier empty: move the split and re-file the far pile
      const unsigned nf = ld_relaxed(far_cnt);
      if (nf == 0) break;
      const int fmin = static_cast<int>(ld_relaxed(far_min));
      const int T_old = T;
      const long long Tn = static_cast<long long>(T > fmin ? T : fmin) + delta;
      T = Tn > kInf ? kInf : static_cast<int>(Tn);
      it++;
      const int nepoch = epoch + 1;
      int* nfar = a.far[nepoch % 3];
      unsigned* nfar_cnt = &ctl->far_cnt[nepoch % 3].v;
      unsigned* nfar_min = &ctl->far_min[nepoch % 3].v;
      if (tid == 0) {
        // pile (epoch + 2) % 3 was consumed by the previous split, one barrier ago
        st_relaxed(&ctl->far_cnt[(epoch + 2) % 3].v, 0u);
        st_relaxed(&ctl->far_min[(epoch + 2) % 3].v, static_cast<unsigned>(kInf));
      }
      se_reset(s_emit);
      for (unsigned base = blockIdx.x * blockDim.x; base < nf; base += stride) {
        const unsigned i = base + threadIdx.x;
        bool to_near = false, keep = false;
        int u = 0, du = 0, udeg = 0;
        if (i < nf) {
          u = ld_relaxed(&far[i]);
          du = ld_relaxed(&a.dist[u]);
          if (du >= T_old) {
            if (du < T) {
              to_near = __hip_atomic_exchange(&a.stamp[u], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != it;
              if (to_near) udeg = degree_of(a.nindex, u);
            } else {
              keep = __hip_atomic_exchange(&a.farstamp[u], nepoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != nepoch;
            }
          }
        }
        frontier_emit_block(s_emit.near, to_near, u, udeg, out, oc);
        se_park(s_emit, keep, u, du, nfar, nfar_cnt);
      }
      se_flush(s_emit, out, oc, nfar, nfar_cnt, nfar_min);
      if (tid == 0) ctl->epochs++;
      epoch = nepoch;
    } else {
      it++;
      se_reset(s_emit);
      auto visit = [&](int v, int e, bool active) {
        const int dv = active ? ld_relaxed(&a.dist[v]) : 0;
        relax(a, dv, e, active, T, it, epoch, far, far_cnt, s_emit, out, oc);
      };
      expand_huge(in, nh, nch, a.nindex, visit);
      expand_mid(in, nm, a.nindex, visit);
      expand_small(in.q, n, a.nindex, visit);
      se_flush(s_emit, out, oc, far, far_cnt, far_min);
    }
    if (!grid_barrier<true>(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      return;
    }
    L++;
  }
  if (tid == 0) ctl->iterations = it;
}

}  // namespace

size_t sssp_workspace_bytes(int nodes)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(SsspControl)) + 13 * Carver::round(v * 4) + 4096;
}

int sssp_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, const int* eweight, int src,
             int* dist)
{
  Carver cv(ctx->ws);
  SsspArgs a{};
  a.ctl = cv.take<SsspControl>(1);
  for (int k = 0; k < 2; k++) {
    a.fq[k].q = cv.take<int>(nodes);
    a.fq[k].mq = cv.take<int>(nodes);
    a.fq[k].hq = cv.take<int>(nodes);
    a.fq[k].hoff = cv.take<unsigned>(nodes);
  }
  for (int k = 0; k < 3; k++) a.far[k] = cv.take<int>(nodes);
  a.stamp = cv.take<int>(nodes);
  a.farstamp = cv.take<int>(nodes);
  a.nodes = nodes;
  a.edges = edges;
  a.src = src;
  a.delta_hint = ctx->hints[INDIGO_HINT_SSSP_DELTA];
  a.nindex = nindex;
  a.nlist = nlist;
  a.eweight = eweight;
  a.dist = dist;
  a.budget = budget_ticks(ctx);
  a.start_ticks = residency_ticks(ctx);
  IG_TRY(zero_async(ctx, a.ctl, sizeof(SsspControl)));
  void* args[] = {&a};
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(sssp_kernel), kSsspBlock, 0, args, 0, nullptr, "sssp_kernel"));
  return INDIGO_OK;
}

int sssp_collect(indigo_ctx* ctx, int* iterations, int* epochs)
{
  Carver cv(ctx->ws);
  SsspControl* ctl = cv.take<SsspControl>(1);
  IG_TRY(ensure_host_pinned(ctx, sizeof(SsspControl)));
  SsspControl* h = static_cast<SsspControl*>(ctx->host_pinned);
  IG_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(SsspControl), hipMemcpyDeviceToHost, ctx->stream));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  IG_TRY(check_abort(ctx, reinterpret_cast<const void*>(sssp_kernel), h->bar.abort.v, "sssp"));
  if (h->status != 0) return set_error(INDIGO_ERR_TIMEOUT, "sssp: persistent kernel stopped (status %d)", h->status);
  if (iterations) *iterations = h->iterations;
  if (epochs) *epochs = h->epochs;
  store_trace(ctx, h->trace);
  return INDIGO_OK;
}

}  // namespace indigo
