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
//     expansion), so stale duplicates only cost work, never correctness;
//   * while the near frontier is small (only short lists, degree sum <=
//     kSsspSmallCap), workgroup 0 runs the near iterations alone with the
//     frontier in LDS and workgroup barriers only (small_near), parking
//     entries at or above the split in the far pile as the grid does, and
//     also the splits while the far pile has at most kSsspSmallCap entries;
//     it hands the frontier back when it outgrows the LDS, or when the
//     frontier empties with a larger (or no) far pile.  Deep sparse graphs spend thousands of iterations on
//     frontiers of tens of vertices, where the grid barrier was most of an
//     iteration's cost.
// delta defaults to a sampled mean weight x 32 / mean degree (at least 1).
#include "sssp.hpp"

#include "frontier.hpp"
#include "trace.hpp"

namespace indigo {

namespace {

constexpr int kSsspBlock = 512;
constexpr int kSsspSmallCap = 2048;  // list entries of a one-workgroup iteration

struct SsspControl {
  GridBarrier bar;
  PaddedU32 small_ack;  // workgroups that have read the counters before a one-workgroup phase
  FrontierCounters cnt[3];
  PaddedU32 far_cnt[3];
  PaddedU32 far_min[3];
  int iterations;
  int epochs;
  int status;
  int delta;
  int small_iters;  // iterations (near and split) run by workgroup 0 alone
  int level_base;   // counter rotation index at which the grid resumes after a one-workgroup phase
  int it_base;      // iteration (stamp) count at which it resumes
  int split_base;   // split T at which it resumes
  int epoch_base;   // far-pile epoch at which it resumes
  int pad[23];
  LevelTrace trace;
};

struct SsspArgs {
  int nodes;
  int edges;
  int src;
  int delta_hint;
  int allow_small;  // the one-workgroup phase may run (INDIGO_HINT_SMALL_FRONTIER != 1)
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

// Mid lists of a near iteration: one wave each, as frontier.hpp's
// expand_mid, or, while there are few of them, k waves each (k up to 16,
// while 2 * k waves per list still fit the grid), wave j of a list taking
// its windows j, j + k, ...  Each window is a chain of device-scope round
// trips (the vertex's distance, the neighbour's check, the atomicMin), so a
// 2000-entry list walked by one wave set a small graph's iteration length.
template <typename F>
IG_DEV void expand_mid_spread(const FrontierQueues& in, unsigned nm, const int* __restrict__ nindex, F&& f)
{
  const int lane = lane_id();
  const unsigned wpb = blockDim.x / kWave;
  const unsigned nwaves = gridDim.x * wpb;
  unsigned k = 1;
  while (k < 16 && static_cast<unsigned long long>(nm) * k * 2 <= nwaves) k *= 2;
  for (unsigned i = blockIdx.x * wpb + threadIdx.x / kWave; i < nm * k; i += nwaves) {
    const int v = ld_relaxed(&in.mq[i / k]);
    const int beg = nindex[v], end = nindex[v + 1];
    for (int base = beg + static_cast<int>(i % k) * kWave; base < end; base += static_cast<int>(k) * kWave) {
      const int e = base + lane;
      f(v, e, e < end);
    }
  }
}

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

// LDS of the one-workgroup phase (shares the kernel's LDS with the emitter,
// which is idle meanwhile).
struct SsspSmallLds {
  int q[kSsspSmallCap];      // near frontier of the current iteration, refilled with the next one
  int incl[kSsspSmallCap];   // inclusive prefix of the frontier's degrees
  int shift[kSsspSmallCap];  // nindex[v] - exclusive prefix: entry k of slot j is nlist[shift[j] + k]
  int dv[kSsspSmallCap];     // distance of slot j's vertex at expansion
  int wsum[kSsspBlock / kWave];
  int cnt;
  int abort_flag;
  unsigned far_min;
};

// Workgroup 0 runs near iterations alone while the frontier's degree sum
// fits the LDS arrays.  Entered at rotation index L with the near frontier in
// the small global queue (n entries); each iteration relaxes every entry as
// relax() does (distance check, atomicMin, then the stamp for an improvement
// below T, which joins the next frontier in LDS, or the far-pile stamp for
// one at or above T, parked in the current epoch's global pile).  When the
// frontier empties and the far pile holds at most kSsspSmallCap entries, it
// also splits as the grid does: T moves to max(T, far minimum) + delta, the
// pile is re-filed (below the new T into the frontier, still above it into
// the next epoch's pile, settled entries dropped) and the piles rotate.  On
// leaving (frontier too large, or empty with a larger or no far pile) it
// re-emits the frontier into the global queues and counters of the rotation
// index reached, cleared first, once every other workgroup has acknowledged
// reading the entry counters (small_ack, as in the BFS kernel's
// one-workgroup phase).  L and it advance by one per iteration and split, as
// in the grid; T and epoch are updated in place.
IG_DEV void small_near(const SsspArgs& a, int& L, int& it, unsigned n, int& T, int& epoch, int delta,
                       SsspControl* ctl, long long deadline, bool* aborted, SsspSmallLds& S)
{
  constexpr int kPer = kSsspSmallCap / kSsspBlock;  // frontier slots per thread
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  int* far = a.far[epoch % 3];
  unsigned* far_cnt = &ctl->far_cnt[epoch % 3].v;
  const int* gin = a.fq[L & 1].q;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) S.q[i] = ld_relaxed(&gin[i]);
  if (threadIdx.x == 0) {
    S.cnt = static_cast<int>(n);
    S.abort_flag = 0;
    S.far_min = static_cast<unsigned>(kInf);
  }
  __syncthreads();
  *aborted = false;
  int cn = S.cnt;
  for (;;) {
    if ((it & 63) == 0 && threadIdx.x == 0) S.abort_flag = kernel_aborted(&ctl->bar, deadline) ? 1 : 0;
    if (cn == 0) {
      // near frontier empty: split here while the far pile is small
      const unsigned nf = ld_relaxed(far_cnt);
      if (nf == 0 || nf > static_cast<unsigned>(kSsspSmallCap)) break;
      const unsigned fm = min(ld_relaxed(&ctl->far_min[epoch % 3].v), S.far_min);
      const int fmin = static_cast<int>(fm);
      trace_level(&ctl->trace, L, nf, 0ull, 3);
      __syncthreads();  // every thread has read S.far_min and the abort flag
      if (S.abort_flag) {
        *aborted = true;
        return;
      }
      const int T_old = T;
      const long long Tn = static_cast<long long>(T > fmin ? T : fmin) + delta;
      T = Tn > kInf ? kInf : static_cast<int>(Tn);
      it++;
      const int nepoch = epoch + 1;
      int* nfar = a.far[nepoch % 3];
      unsigned* nfar_cnt = &ctl->far_cnt[nepoch % 3].v;
      if (threadIdx.x == 0) {
        // pile (epoch + 2) % 3 was consumed by the previous split
        st_relaxed(&ctl->far_cnt[(epoch + 2) % 3].v, 0u);
        st_relaxed(&ctl->far_min[(epoch + 2) % 3].v, static_cast<unsigned>(kInf));
        S.far_min = static_cast<unsigned>(kInf);
        S.cnt = 0;
        ctl->epochs++;
      }
      __syncthreads();
      for (unsigned base = 0; base < nf; base += blockDim.x) {
        const unsigned i = base + threadIdx.x;
        bool to_near = false, keep = false;
        int u = 0, du = 0;
        if (i < nf) {
          u = ld_relaxed(&far[i]);
          du = ld_relaxed(&a.dist[u]);
          if (du >= T_old) {
            if (du < T)
              to_near = __hip_atomic_exchange(&a.stamp[u], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != it;
            else
              keep = __hip_atomic_exchange(&a.farstamp[u], nepoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                     nepoch;
          }
        }
        const unsigned long long m = __ballot(to_near);
        if (m != 0) {
          const int leader = __ffsll(static_cast<long long>(m)) - 1;
          int b = 0;
          if (lane == leader) b = atomicAdd(&S.cnt, __popcll(m));
          b = wave_shfl(b, leader);
          // at most nf <= kSsspSmallCap entries in all
          if (to_near) S.q[b + lanes_below(m)] = u;
        }
        wave_append(keep, u, nfar, nfar_cnt);
        const int km = wave_min(keep ? du : kInf);
        if (lane == 0 && km != kInf) atomicMin(&S.far_min, static_cast<unsigned>(km));
      }
      epoch = nepoch;
      far = nfar;
      far_cnt = nfar_cnt;
      __syncthreads();
      cn = S.cnt;
      L++;
      __syncthreads();
      continue;
    }
    // round 1: index pairs and distances of this thread's run of slots
    const int i0 = min(cn, static_cast<int>(threadIdx.x) * kPer);
    const int i1 = min(cn, i0 + kPer);
    int beg[kPer], deg[kPer], dvs[kPer];
    int mine = 0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const bool ok = i0 + j < i1;
      const int v = ok ? S.q[i0 + j] : 0;
      beg[j] = ok ? a.nindex[v] : 0;
      deg[j] = ok ? a.nindex[v + 1] - beg[j] : 0;
      dvs[j] = ok ? ld_relaxed(&a.dist[v]) : 0;
      mine += deg[j];
    }
    const int winc = wave_inclusive_scan(mine);
    if (lane == kWave - 1) S.wsum[wid] = winc;
    __syncthreads();
    if (S.abort_flag) {
      *aborted = true;
      return;
    }
    int before = 0, total = 0;
    for (int w = 0; w < kSsspBlock / kWave; w++) {
      before += w < wid ? S.wsum[w] : 0;
      total += S.wsum[w];
    }
    if (total > kSsspSmallCap) break;  // S.q still holds this iteration's frontier
    trace_level(&ctl->trace, L, static_cast<unsigned>(cn), static_cast<unsigned long long>(total), 2);
    int run = before + winc - mine;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      if (i0 + j < i1) {
        S.incl[i0 + j] = run + deg[j];
        S.shift[i0 + j] = beg[j] - run;
        S.dv[i0 + j] = dvs[j];
        run += deg[j];
      }
    }
    if (threadIdx.x == 0) S.cnt = 0;
    __syncthreads();
    it++;
    // round 2: the list entries, spread evenly over the workgroup
    int u[kPer], nd[kPer];
    bool act[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const int k = j * kSsspBlock + static_cast<int>(threadIdx.x);
      act[j] = k < total;
      u[j] = 0;
      nd[j] = kInf;
      if (act[j]) {
        int lo = 0, hi = cn - 1;  // owner: the first slot with incl > k
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (S.incl[mid] > k) hi = mid;
          else lo = mid + 1;
        }
        const int e = S.shift[lo] + k;
        u[j] = a.nlist[e];
        nd[j] = S.dv[lo] + a.eweight[e];
      }
    }
    // round 3: relaxations (device-scope reads, as relax())
    bool imp[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) imp[j] = act[j] && nd[j] < ld_relaxed(&a.dist[u[j]]);
#pragma unroll
    for (int j = 0; j < kPer; j++) imp[j] = imp[j] && nd[j] < atomic_min_relaxed(&a.dist[u[j]], nd[j]);
    bool near[kPer], park[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      near[j] = imp[j] && nd[j] < T &&
                __hip_atomic_exchange(&a.stamp[u[j]], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != it;
      park[j] = imp[j] && nd[j] >= T &&
                __hip_atomic_exchange(&a.farstamp[u[j]], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
      const unsigned long long m = __ballot(near[j]);
      if (m != 0) {
        const int leader = __ffsll(static_cast<long long>(m)) - 1;
        int b = 0;
        if (lane == leader) b = atomicAdd(&S.cnt, __popcll(m));
        b = wave_shfl(b, leader);
        // stamps admit each vertex once per iteration: at most total <= kSsspSmallCap
        if (near[j]) S.q[b + lanes_below(m)] = u[j];
      }
      wave_append(park[j], u[j], far, far_cnt);
      const int pm = wave_min(park[j] ? nd[j] : kInf);
      if (lane == 0 && pm != kInf) atomicMin(&S.far_min, static_cast<unsigned>(pm));
    }
    __syncthreads();
    cn = S.cnt;
    L++;
    __syncthreads();
  }
  // hand the frontier back to the grid, once all others have read the entry
  // iteration's counters
  if (threadIdx.x == 0) {
    if (S.far_min != static_cast<unsigned>(kInf)) atomic_min_relaxed(&ctl->far_min[epoch % 3].v, S.far_min);
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
    return;
  }
  for (int base = 0; base < cn; base += blockDim.x) {
    const int i = base + static_cast<int>(threadIdx.x);
    const int v = i < cn ? S.q[i] : 0;
    const int d = i < cn ? degree_of(a.nindex, v) : 0;
    frontier_emit(d > 0, v, d, a.fq[L & 1], &ctl->cnt[L % 3]);
  }
}

// 6 waves per SIMD: three 512-thread workgroups per CU, the persistent grid's
// residency (at most 80 VGPRs)
__global__ __launch_bounds__(kSsspBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void sssp_kernel(SsspArgs a)
{
  __shared__ union {
    SsspEmit emit;
    SsspSmallLds small;
  } s_lds;
  SsspEmit& s_emit = s_lds.emit;
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
    const unsigned long long m = ld_relaxed(&ic->m);
    const bool small =
        a.allow_small && !split && nm == 0 && nh == 0 && m <= static_cast<unsigned long long>(kSsspSmallCap);
    trace_level(&ctl->trace, L, split ? ld_relaxed(far_cnt) : n + nm + nh, split ? 0ull : m,
                split ? 3 : (small ? 2 : 0));
    if (small) {
      // one-workgroup phase; the other workgroups acknowledge that they have
      // read this iteration's counters and wait at the next barrier
      if (blockIdx.x != 0 && threadIdx.x == 0)
        __hip_atomic_fetch_add(&ctl->small_ack.v, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (blockIdx.x == 0) {
        bool aborted = false;
        int L2 = L, it2 = it, T2 = T, epoch2 = epoch;
        small_near(a, L2, it2, n, T2, epoch2, delta, ctl, deadline, &aborted, s_lds.small);
        if (threadIdx.x == 0) {
          st_relaxed(&ctl->level_base, L2);
          st_relaxed(&ctl->it_base, it2);
          st_relaxed(&ctl->split_base, T2);
          st_relaxed(&ctl->epoch_base, epoch2);
          ctl->small_iters += L2 - L;
        }
      }
      if (!grid_barrier<true>(&ctl->bar, deadline)) {
        if (tid == 0) ctl->status = 1;
        return;
      }
      const int L2 = ld_relaxed(&ctl->level_base);
      if (L2 == L) {  // cannot happen: the phase always runs at least one iteration
        if (tid == 0) ctl->status = 2;
        return;
      }
      L = L2;
      it = ld_relaxed(&ctl->it_base);
      T = ld_relaxed(&ctl->split_base);
      epoch = ld_relaxed(&ctl->epoch_base);
      continue;
    }
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
      expand_mid_spread(in, nm, a.nindex, visit);
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
  a.allow_small = ctx->hints[INDIGO_HINT_SMALL_FRONTIER] != 1;
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
