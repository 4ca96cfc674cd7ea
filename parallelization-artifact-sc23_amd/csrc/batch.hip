// Batched engine: many graphs and many sources per graph in one persistent
// launch (MI355X, gfx950).
//
// Every query becomes a task: multi-source BFS (up to 64 sources), multi-
// source SSSP (up to 32), CC or MIS on one graph.  The host sorts tasks by
// size (largest first) into a device-side queue; one persistent kernel,
// sized to co-residency and checked by a residency barrier, serves it.
//
// Gangs.  A task runs on a gang of k workgroups, k = the task's CSR entries
// / 2^18 rounded up to a power of two (1..64): small graphs take one
// workgroup and synchronise with __syncthreads only, large ones a gang with
// its own bounded barrier.  A free workgroup joins the first task in queue
// order that still has a free slot (one atomic on the task's `joined`
// counter); the gang starts when all k members have arrived.  Because every
// workgroup is resident and gangs are filled strictly in queue order, a gang
// only ever waits for workgroups that are busy with earlier tasks, so
// formation cannot deadlock.
//
// Memory discipline (as the coherent-mode persistent kernels): data one
// thread writes and another reads within a task goes through device-scope
// relaxed loads / stores / atomics, and each barrier first drains the
// wave's memory operations; barriers are relaxed counters.
//
// Algorithms (same results as the single-query kernels and the reference):
//   BFS : level-synchronous pull over 64-bit masks: frontier[v] holds the
//         sources whose BFS reached v at the previous level; a vertex ORs
//         its neighbours' frontier masks, keeps the bits not yet seen and
//         writes the level into the slab rows of those sources.
//   SSSP: Bellman-Ford over a per-source distance slab with per-vertex
//         32-bit active masks (triple-buffered); relaxations by atomicMin,
//         a vertex is re-activated for a source when its distance drops.
//   CC  : union-find hooking the larger root under the smaller (labels =
//         component minima), then flattening.
//   MIS : the greedy fixed point in descending (hash(v + 712313887), v)
//         order, two phases per round (decide, then exclude).
//   PR  : the Jacobi recurrence of the single-query kernel (FP32 scores,
//         Neumaier-compensated gathers: one lane per short list in entry
//         order, one wave per list of kPrWaveList or more), stopping when the
//         FP64 L1 change is <= 1e-4 or after 100 sweeps; every gang member
//         adds the members' FP64 partials in the same order, so all take the
//         same decision.
// Lists are walked 64 entries per wave step with the in-wave load balancer,
// so skewed degrees do not serialise a lane.
#include "batch.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "csum.hpp"
#include "device_common.hpp"

namespace indigo {

namespace {

constexpr int kBatchBlock = 512;
constexpr int kBatchWaves = kBatchBlock / kWave;
constexpr int kMaxGang = 64;
constexpr int kGangShift = 18;  // CSR entries per gang member
constexpr int kPrWaveList = 32;  // PR: lists walked by a whole wave
constexpr float kDamp = 0.85f;

struct alignas(128) GangState {
  unsigned joined;   // members that took a slot (grows past the gang size)
  unsigned arrive;   // barrier arrivals of the current generation
  unsigned gen;      // barrier generation
  unsigned flag[3];  // per-round "something changed" flags, triple-buffered
  unsigned pad[26];
};

struct BatchTask {
  int alg, gang, nodes, edges;
  const int* nindex;
  const int* nlist;
  const int* eweight;
  int* out;
  void* scratch;  // BFS: 3 x u64[nodes]; SSSP: 3 x u32[nodes]; PR: 2 x f32[nodes] + 2 x f64[kMaxGang]
  int nsrc;
  int src[INDIGO_MAX_BFS_SOURCES];
};

struct alignas(128) BatchControl {
  GridBarrier bar;
  unsigned head;  // first task that may still have a free slot
  unsigned pad[31];
};

struct Team {
  int rank, size;
  GangState* gs;
  GridBarrier* gb;
  long long deadline;
};

// Gang barrier (bounded).  Size 1: the workgroup barrier suffices.
IG_DEV bool team_sync(const Team& tm)
{
  __shared__ int s_ok;
  wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    if (tm.size > 1) {
      const unsigned gen0 = ld_relaxed(&tm.gs->gen);
      wave_drain();
      const unsigned a = atomic_add_relaxed(&tm.gs->arrive, 1u);
      if (a == static_cast<unsigned>(tm.size) - 1u) {
        st_relaxed(&tm.gs->arrive, 0u);
        wave_drain();
        st_relaxed(&tm.gs->gen, gen0 + 1u);
      } else {
        int backoff = 1;
        while (ld_relaxed(&tm.gs->gen) == gen0) {
          if (ld_relaxed(&tm.gb->abort.v) != 0u) break;
          if (now_ticks() > tm.deadline) {
            unsigned expected = 0u;
            __hip_atomic_compare_exchange_strong(&tm.gb->abort.v, &expected, kAbortTimeout, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          for (int k = 0; k < backoff; k++) __builtin_amdgcn_s_sleep(2);
          backoff = backoff < 64 ? backoff * 2 : 64;
        }
      }
    } else if (now_ticks() > tm.deadline) {
      unsigned expected = 0u;
      __hip_atomic_compare_exchange_strong(&tm.gb->abort.v, &expected, kAbortTimeout, __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (ld_relaxed(&tm.gb->abort.v) != 0u) ok = 0;
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

IG_DEV void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-collective: walks the CSR entries of the lanes' vertices v0 + lane
// (those with `active`), 64 per step; f(o, e, valid) gets the owner lane o
// and the entry e, with all lanes converged.
template <typename F>
IG_DEV void wave_entries(int v0, int nodes, const int* __restrict__ nindex, bool active, F&& f)
{
  const int lane = lane_id();
  const int v = v0 + lane;
  const int b = active && v < nodes ? nindex[v] : 0;
  const int len = active && v < nodes ? nindex[v + 1] - b : 0;
  const int incl = wave_inclusive_scan(len);
  const int total = wave_bcast(incl, kWave - 1);
  const int shift = b - (incl - len);
  for (int base = 0; base < total; base += kWave) {
    const int k = base + lane;
    const int o = min(wave_owner(incl, k), kWave - 1);
    const int e = wave_shfl(shift, o) + k;
    f(o, e, k < total);
  }
}

// The team's share of 64-vertex chunks: chunk c for c = first, first + step, ...
struct Chunks {
  int first, step, count;
};

IG_DEV Chunks team_chunks(const Team& tm, int nodes)
{
  return Chunks{tm.rank * kBatchWaves + static_cast<int>(threadIdx.x) / kWave, tm.size * kBatchWaves,
                (nodes + kWave - 1) / kWave};
}

IG_DEV void set_flag(unsigned* flag, bool any)
{
  if (__ballot(any) != 0ull && lane_id() == 0) st_relaxed(flag, 1u);
}

// ---- multi-source BFS -------------------------------------------------------
IG_DEV bool run_bfs(const BatchTask& T, const Team& tm, unsigned long long (*s_acc)[kWave])
{
  const int V = T.nodes;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  unsigned long long* seen = static_cast<unsigned long long*>(T.scratch);
  unsigned long long* fr[2] = {seen + V, seen + 2 * static_cast<size_t>(V)};
  const unsigned long long all = T.nsrc >= 64 ? ~0ull : ((1ull << T.nsrc) - 1ull);
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) {
      st_relaxed(&seen[v], 0ull);
      st_relaxed(&fr[0][v], 0ull);
      st_relaxed(&fr[1][v], 0ull);
    }
  }
  if (!team_sync(tm)) return false;
  if (tm.rank == 0 && static_cast<int>(threadIdx.x) < T.nsrc) {
    const int i = threadIdx.x;
    const int s = T.src[i];
    __hip_atomic_fetch_or(&seen[s], 1ull << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_or(&fr[0][s], 1ull << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st_relaxed(&T.out[static_cast<size_t>(i) * V + s], 0);
  }
  if (!team_sync(tm)) return false;
  for (int lvl = 1;; lvl++) {
    const unsigned long long* cur = fr[(lvl - 1) & 1];
    unsigned long long* nxt = fr[lvl & 1];
    unsigned* flag = &tm.gs->flag[lvl % 3];
    if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&tm.gs->flag[(lvl + 1) % 3], 0u);
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const unsigned long long sv = v < V ? ld_relaxed(&seen[v]) : all;
      const unsigned long long missing = all & ~sv;
      s_acc[wid][lane] = 0ull;
      wave_lds_sync();
      wave_entries(c * kWave, V, T.nindex, missing != 0ull, [&](int o, int e, bool valid) {
        if (valid) {
          const unsigned long long m = ld_relaxed(&cur[T.nlist[e]]);
          if (m != 0ull) atomicOr(&s_acc[wid][o], m);
        }
      });
      wave_lds_sync();
      const unsigned long long nb = s_acc[wid][lane] & missing;
      if (v < V) {
        st_relaxed(&nxt[v], nb);
        if (nb != 0ull) {
          st_relaxed(&seen[v], sv | nb);
          for (unsigned long long m = nb; m != 0ull; m &= m - 1ull)
            T.out[static_cast<size_t>(__ffsll(static_cast<long long>(m)) - 1) * V + v] = lvl;
        }
      }
      set_flag(flag, nb != 0ull);
    }
    if (!team_sync(tm)) return false;
    if (ld_relaxed(flag) == 0u) break;
  }
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v >= V) continue;
    for (unsigned long long m = all & ~ld_relaxed(&seen[v]); m != 0ull; m &= m - 1ull)
      T.out[static_cast<size_t>(__ffsll(static_cast<long long>(m)) - 1) * V + v] = INT_MAX;
  }
  return true;
}

// ---- multi-source SSSP -------------------------------------------------------
IG_DEV bool run_sssp(const BatchTask& T, const Team& tm)
{
  const int V = T.nodes;
  const int lane = lane_id();
  unsigned* act[3] = {static_cast<unsigned*>(T.scratch), static_cast<unsigned*>(T.scratch) + V,
                      static_cast<unsigned*>(T.scratch) + 2 * static_cast<size_t>(V)};
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v >= V) continue;
    st_relaxed(&act[0][v], 0u);
    st_relaxed(&act[1][v], 0u);
    st_relaxed(&act[2][v], 0u);
    for (int i = 0; i < T.nsrc; i++) st_relaxed(&T.out[static_cast<size_t>(i) * V + v], INT_MAX);
  }
  if (!team_sync(tm)) return false;
  if (tm.rank == 0 && static_cast<int>(threadIdx.x) < T.nsrc) {
    const int i = threadIdx.x;
    const int s = T.src[i];
    st_relaxed(&T.out[static_cast<size_t>(i) * V + s], 0);
    __hip_atomic_fetch_or(&act[0][s], 1u << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!team_sync(tm)) return false;
  for (int r = 0;; r++) {
    const unsigned* cur = act[r % 3];
    unsigned* nxt = act[(r + 1) % 3];
    unsigned* clr = act[(r + 2) % 3];
    unsigned* flag = &tm.gs->flag[r % 3];
    if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&tm.gs->flag[(r + 1) % 3], 0u);
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const unsigned m = v < V ? ld_relaxed(&cur[v]) : 0u;
      if (v < V) st_relaxed(&clr[v], 0u);
      bool changed = false;
      wave_entries(c * kWave, V, T.nindex, m != 0u, [&](int o, int e, bool valid) {
        const unsigned mo = static_cast<unsigned>(wave_shfl(static_cast<int>(m), o));
        if (!valid) return;
        const int ov = c * kWave + o;
        const int u = T.nlist[e];
        const int w = T.eweight[e];
        for (unsigned bits = mo; bits != 0u; bits &= bits - 1u) {
          const int i = __ffs(static_cast<int>(bits)) - 1;
          const size_t row = static_cast<size_t>(i) * V;
          const int nd = ld_relaxed(&T.out[row + ov]) + w;
          if (nd < ld_relaxed(&T.out[row + u])) {
            const int old = atomic_min_relaxed(&T.out[row + u], nd);
            if (nd < old) {
              __hip_atomic_fetch_or(&nxt[u], 1u << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              changed = true;
            }
          }
        }
      });
      set_flag(flag, changed);
    }
    if (!team_sync(tm)) return false;
    if (ld_relaxed(flag) == 0u) break;
  }
  return true;
}

// ---- CC ------------------------------------------------------------------------
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
    ra = find_root(parent, ra == hi ? expected : ra);
    rb = find_root(parent, rb == hi ? expected : rb);
  }
}

IG_DEV bool run_cc(const BatchTask& T, const Team& tm)
{
  const int V = T.nodes;
  const int lane = lane_id();
  int* parent = T.out;
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) st_relaxed(&parent[v], v);
  }
  if (!team_sync(tm)) return false;
  for (int c = ch.first; c < ch.count; c += ch.step) {
    wave_entries(c * kWave, V, T.nindex, true, [&](int o, int e, bool valid) {
      if (!valid) return;
      const int v = c * kWave + o;
      const int u = T.nlist[e];
      if (u < v) link(parent, v, u);
    });
  }
  if (!team_sync(tm)) return false;
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) st_relaxed(&parent[v], find_root(parent, v));
  }
  return true;
}

// ---- MIS -------------------------------------------------------------------------
IG_DEV unsigned long long mis_key(int v)
{
  return (static_cast<unsigned long long>(mis_hash(static_cast<unsigned>(v) + 712313887u)) << 32) |
         static_cast<unsigned>(v);
}

IG_DEV bool run_mis(const BatchTask& T, const Team& tm, int (*s_flag)[kWave])
{
  const int V = T.nodes;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  int* status = T.out;
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) st_relaxed(&status[v], 0);
  }
  if (!team_sync(tm)) return false;
  for (int r = 0;; r++) {
    unsigned* flag = &tm.gs->flag[r % 3];
    if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&tm.gs->flag[(r + 1) % 3], 0u);
    // decide: an undecided vertex joins when no neighbour is included and no
    // undecided neighbour has a larger key
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const bool und = v < V && ld_relaxed(&status[v]) == 0;
      s_flag[wid][lane] = 0;
      wave_lds_sync();
      wave_entries(c * kWave, V, T.nindex, und, [&](int o, int e, bool valid) {
        if (!valid) return;
        const int ov = c * kWave + o;
        const int u = T.nlist[e];
        const int su = ld_relaxed(&status[u]);
        if (su == 1 || (su == 0 && u != ov && mis_key(u) > mis_key(ov))) s_flag[wid][o] = 1;
      });
      wave_lds_sync();
      if (und && s_flag[wid][lane] == 0) st_relaxed(&status[v], 1);
    }
    if (!team_sync(tm)) return false;
    // exclude the undecided neighbours of included vertices
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const bool und = v < V && ld_relaxed(&status[v]) == 0;
      s_flag[wid][lane] = 0;
      wave_lds_sync();
      wave_entries(c * kWave, V, T.nindex, und, [&](int o, int e, bool valid) {
        if (valid && ld_relaxed(&status[T.nlist[e]]) == 1) s_flag[wid][o] = 1;
      });
      wave_lds_sync();
      const bool excl = und && s_flag[wid][lane] != 0;
      if (excl) st_relaxed(&status[v], 2);
      set_flag(flag, und && !excl);
    }
    if (!team_sync(tm)) return false;
    if (ld_relaxed(flag) == 0u) break;
  }
  return true;
}

// ---- PageRank ----------------------------------------------------------------------
IG_DEV bool run_pr(const BatchTask& T, const Team& tm, double* s_red)
{
  const int V = T.nodes;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  float* scores = reinterpret_cast<float*>(T.out);
  float* contrib[2] = {static_cast<float*>(T.scratch), static_cast<float*>(T.scratch) + V};
  double* slots = reinterpret_cast<double*>(static_cast<float*>(T.scratch) + 2 * static_cast<size_t>(V) +
                                            ((2 * static_cast<size_t>(V)) & 1));  // 8-byte aligned
  const float inv_n = 1.0f / static_cast<float>(V);
  const float base = (1.0f - kDamp) / static_cast<float>(V);
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v >= V) continue;
    scores[v] = inv_n;
    st_relaxed(&contrib[0][v], inv_n / static_cast<float>(T.nindex[v + 1] - T.nindex[v]));
  }
  if (!team_sync(tm)) return false;
  int it = 0;
  for (;;) {
    it++;
    const float* cin = contrib[(it - 1) & 1];
    float* cout = contrib[it & 1];
    double d = 0.0;
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const int beg = v < V ? T.nindex[v] : 0;
      const int deg = v < V ? T.nindex[v + 1] - beg : 0;
      auto finish = [&](int x, int xdeg, float total) {
        const float old = scores[x];
        const float s = base + kDamp * total;
        scores[x] = s;
        st_relaxed(&cout[x], s / static_cast<float>(xdeg));
        return fabs(static_cast<double>(s) - static_cast<double>(old));
      };
      if (v < V && deg < kPrWaveList) {
        CSum acc;
        for (int e = beg; e < beg + deg; e++) acc.add(ld_relaxed(&cin[T.nlist[e]]));
        d += finish(v, deg, acc.value());
      }
      for (unsigned long long m = __ballot(v < V && deg >= kPrWaveList); m != 0ull; m &= m - 1ull) {
        const int o = __ffsll(static_cast<long long>(m)) - 1;
        const int x = c * kWave + o;
        const int xb = wave_bcast(beg, o), xd = wave_bcast(deg, o);
        CSum acc;
        for (int e = xb + lane; e < xb + xd; e += kWave) acc.add(ld_relaxed(&cin[T.nlist[e]]));
        acc = wave_merge(acc);
        if (lane == 0) d += finish(x, xd, acc.value());
      }
    }
    // the member's FP64 partial (fixed order: wave tree, then waves in order)
    d = wave_sum(d);
    if (lane == 0) s_red[wid] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kBatchWaves; w++) t += s_red[w];
      st_relaxed(&slots[(it & 1) * kMaxGang + tm.rank], t);
    }
    if (!team_sync(tm)) return false;
    double total = 0.0;
    for (int r = 0; r < tm.size; r++) total += ld_relaxed(&slots[(it & 1) * kMaxGang + r]);
    if (!(total > static_cast<double>(0.0001f) && it < 100)) break;
  }
  if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&T.out[V], it);
  return true;
}

__global__ __launch_bounds__(kBatchBlock) void batch_kernel(const BatchTask* __restrict__ tasks, int ntasks,
                                                            GangState* gangs, BatchControl* ctl, long long budget,
                                                            long long residency_ticks)
{
  __shared__ int s_task, s_rank;
  __shared__ unsigned long long s_acc[kBatchWaves][kWave];
  __shared__ int s_flag[kBatchWaves][kWave];
  __shared__ double s_red[kBatchWaves];
  if (!residency_barrier(&ctl->bar, residency_ticks)) return;
  const long long deadline = now_ticks() + budget;
  for (;;) {
    if (threadIdx.x == 0) {
      int t = static_cast<int>(ld_relaxed(&ctl->head));
      int rank = -1;
      while (t < ntasks) {
        const unsigned r = atomic_add_relaxed(&gangs[t].joined, 1u);
        if (r < static_cast<unsigned>(min(tasks[t].gang, static_cast<int>(gridDim.x)))) {
          rank = static_cast<int>(r);
          break;
        }
        atomic_max_relaxed(&ctl->head, static_cast<unsigned>(t + 1));
        t++;
      }
      s_task = t;
      s_rank = rank;
    }
    __syncthreads();
    const int t = s_task;
    const int rank = s_rank;
    __syncthreads();
    if (t >= ntasks) break;
    const BatchTask& T = tasks[t];
    // a relaunch after a failed residency check may run a smaller grid
    const Team tm{rank, min(T.gang, static_cast<int>(gridDim.x)), &gangs[t], &ctl->bar, deadline};
    if (!team_sync(tm)) break;  // the gang assembles
    bool ok = true;
    if (T.alg == INDIGO_QUERY_BFS) ok = run_bfs(T, tm, s_acc);
    else if (T.alg == INDIGO_QUERY_SSSP) ok = run_sssp(T, tm);
    else if (T.alg == INDIGO_QUERY_CC) ok = run_cc(T, tm);
    else if (T.alg == INDIGO_QUERY_PR) ok = run_pr(T, tm, s_red);
    else ok = run_mis(T, tm, s_flag);
    if (!ok) break;
  }
}

size_t scratch_bytes(int alg, int nodes)
{
  const size_t v = static_cast<size_t>(nodes);
  if (alg == INDIGO_QUERY_BFS) return Carver::round(3 * v * 8);
  if (alg == INDIGO_QUERY_SSSP) return Carver::round(3 * v * 4);
  if (alg == INDIGO_QUERY_PR) return Carver::round(2 * v * 4 + 8 + 2 * kMaxGang * 8);
  return 0;
}

int gang_for(long long entries, int grid)
{
  int k = 1;
  while (k < kMaxGang && (static_cast<long long>(k) << kGangShift) < entries) k <<= 1;
  return std::min(k, std::max(1, grid));
}

}  // namespace

// (Re)launches the batch recorded in ctx->pending_batch; records the stop
// event without waiting.
int batch_grid_limit(indigo_ctx* ctx, int* limit)
{
  int per_cu = 0;
  IG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(batch_kernel),
                                                         kBatchBlock, 0));
  *limit = std::max(1, per_cu * ctx->compute_units);
  return INDIGO_OK;
}

static int launch_pending(indigo_ctx* ctx)
{
  BatchTask* d_tasks = static_cast<BatchTask*>(ctx->pending_batch.tasks);
  GangState* gangs = static_cast<GangState*>(ctx->pending_batch.gangs);
  BatchControl* ctl = static_cast<BatchControl*>(ctx->pending_batch.ctl);
  int ntasks = ctx->pending_batch.ntasks;
  long long budget = budget_ticks(ctx);
  long long res_ticks = residency_ticks(ctx);
  IG_TRY(timing_begin(ctx));
  IG_TRY(zero_async(ctx, ctl, sizeof(BatchControl)));
  IG_TRY(zero_async(ctx, gangs, static_cast<size_t>(ntasks) * sizeof(GangState)));
  void* args[] = {&d_tasks, &ntasks, &gangs, &ctl, &budget, &res_ticks};
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(batch_kernel), kBatchBlock, 0, args,
                     ctx->pending_batch.grid_cap, nullptr, "batch_kernel"));
  return timing_record_stop(ctx);
}

int batch_launch(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
                 int nqueries, int max_workgroups)
{
  if (ctx->pending_batch.active) return set_error(INDIGO_ERR_INVALID_ARG, "batch: the context has a batch in flight");
  if (nqueries < 0 || ngraphs < 0 || (nqueries > 0 && (graphs == nullptr || queries == nullptr)))
    return set_error(INDIGO_ERR_INVALID_ARG, "batch: bad graph / query table");
  if (max_workgroups < 0) return set_error(INDIGO_ERR_INVALID_ARG, "batch: max_workgroups < 0");
  if (nqueries == 0) {
    // an empty batch completes at once: record a stop event for batch_ready
    IG_TRY(timing_begin(ctx));
    IG_TRY(timing_record_stop(ctx));
    ctx->pending_batch = {};
    ctx->pending_batch.active = true;
    return INDIGO_OK;
  }
  for (int g = 0; g < ngraphs; g++) {
    const indigo_graph_ref& G = graphs[g];
    IG_TRY(check_graph_args(G.nodes, G.edges, G.d_nindex, G.d_nlist));
  }
  for (int q = 0; q < nqueries; q++) {
    const indigo_query& Q = queries[q];
    if (Q.graph < 0 || Q.graph >= ngraphs) return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d graph out of range", q);
    const indigo_graph_ref& G = graphs[Q.graph];
    if (Q.d_out == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d output is NULL", q);
    if (Q.alg == INDIGO_QUERY_BFS || Q.alg == INDIGO_QUERY_SSSP) {
      const int cap = Q.alg == INDIGO_QUERY_BFS ? INDIGO_MAX_BFS_SOURCES : INDIGO_MAX_SSSP_SOURCES;
      if (Q.nsrc < 1 || Q.nsrc > cap)
        return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d needs 1..%d sources (got %d)", q, cap, Q.nsrc);
      for (int i = 0; i < Q.nsrc; i++)
        if (Q.src[i] < 0 || Q.src[i] >= G.nodes)
          return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d source %d out of range", q, Q.src[i]);
      if (Q.alg == INDIGO_QUERY_SSSP && G.d_eweight == nullptr && G.edges > 0)
        return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d is SSSP on a graph without weights", q);
    } else if (Q.alg != INDIGO_QUERY_CC && Q.alg != INDIGO_QUERY_MIS && Q.alg != INDIGO_QUERY_PR) {
      return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d has unknown algorithm %d", q, Q.alg);
    }
  }
  // grid: occupancy x CUs (capped), learned residency caps as for the other persistent kernels
  int per_cu = 0;
  IG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(batch_kernel),
                                                         kBatchBlock, 0));
  const int grid_cap = std::max(1, max_workgroups > 0 ? std::min(max_workgroups, per_cu * ctx->compute_units)
                                                      : per_cu * ctx->compute_units);
  // tasks, largest gang and work first
  std::vector<int> order(static_cast<size_t>(nqueries));
  std::vector<long long> work(static_cast<size_t>(nqueries));
  for (int q = 0; q < nqueries; q++) {
    order[q] = q;
    const indigo_graph_ref& G = graphs[queries[q].graph];
    work[q] = static_cast<long long>(G.edges) + G.nodes;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  size_t scratch_total = 0;
  for (int q = 0; q < nqueries; q++) scratch_total += scratch_bytes(queries[q].alg, graphs[queries[q].graph].nodes);
  const size_t n = static_cast<size_t>(nqueries);
  const size_t bytes = Carver::round(sizeof(BatchControl)) + Carver::round(n * sizeof(GangState)) +
                       Carver::round(n * sizeof(BatchTask)) + scratch_total + 4096;
  IG_TRY(ensure_workspace(ctx, bytes));
  IG_TRY(ensure_host_pinned(ctx, n * sizeof(BatchTask)));
  Carver cv(ctx->ws);
  BatchControl* ctl = cv.take<BatchControl>(1);
  GangState* gangs = cv.take<GangState>(n);
  BatchTask* d_tasks = cv.take<BatchTask>(n);
  unsigned char* scratch = cv.take<unsigned char>(scratch_total + 256);
  BatchTask* h_tasks = static_cast<BatchTask*>(ctx->host_pinned);
  size_t off = 0;
  for (size_t i = 0; i < n; i++) {
    const indigo_query& Q = queries[order[i]];
    const indigo_graph_ref& G = graphs[Q.graph];
    BatchTask& T = h_tasks[i];
    std::memset(&T, 0, sizeof(T));
    T.alg = Q.alg;
    T.gang = gang_for(work[order[i]], grid_cap);
    T.nodes = G.nodes;
    T.edges = G.edges;
    T.nindex = G.d_nindex;
    T.nlist = G.d_nlist;
    T.eweight = G.d_eweight;
    T.out = Q.d_out;
    T.scratch = scratch + off;
    off += scratch_bytes(Q.alg, G.nodes);
    T.nsrc = (Q.alg == INDIGO_QUERY_BFS || Q.alg == INDIGO_QUERY_SSSP) ? Q.nsrc : 0;
    for (int s = 0; s < T.nsrc; s++) T.src[s] = Q.src[s];
  }
  IG_HIP_TRY(hipMemcpyAsync(d_tasks, h_tasks, n * sizeof(BatchTask), hipMemcpyHostToDevice, ctx->stream));
  ctx->pending_batch.active = true;
  ctx->pending_batch.tasks = d_tasks;
  ctx->pending_batch.gangs = gangs;
  ctx->pending_batch.ctl = ctl;
  ctx->pending_batch.ntasks = nqueries;
  ctx->pending_batch.grid_cap = grid_cap;
  return launch_pending(ctx);
}

int batch_ready(indigo_ctx* ctx, int* ready)
{
  if (!ctx->pending_batch.active) return set_error(INDIGO_ERR_INVALID_ARG, "batch: nothing in flight");
  const hipError_t e = hipEventQuery(ctx->ev_stop);
  if (e == hipErrorNotReady) {
    *ready = 0;
    return INDIGO_OK;
  }
  IG_HIP_TRY(e);
  *ready = 1;
  return INDIGO_OK;
}

int batch_wait(indigo_ctx* ctx)
{
  if (!ctx->pending_batch.active) return set_error(INDIGO_ERR_INVALID_ARG, "batch: nothing in flight");
  for (;;) {
    const int ntasks = ctx->pending_batch.ntasks;
    const int rc0 = timing_finish(ctx, 3, ntasks);
    if (rc0 != INDIGO_OK || ntasks == 0) {
      ctx->pending_batch.active = false;
      return rc0;
    }
    BatchControl* ctl = static_cast<BatchControl*>(ctx->pending_batch.ctl);
    IG_HIP_TRY(hipMemcpyAsync(ctx->host_pinned, &ctl->bar.abort.v, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
    IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const unsigned abort_code = *static_cast<unsigned*>(ctx->host_pinned);
    const int rc = check_abort(ctx, reinterpret_cast<const void*>(batch_kernel), abort_code, "batch");
    if (rc != kNotResident) {
      ctx->pending_batch.active = false;
      return rc;
    }
    // the grid did not assemble: relaunch with the lowered residency cap (the
    // device copy of the task table is intact)
    const int lrc = launch_pending(ctx);
    if (lrc != INDIGO_OK) {
      ctx->pending_batch.active = false;
      return lrc;
    }
  }
}

int batch_run(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
              int nqueries, int max_workgroups)
{
  IG_TRY(batch_launch(ctx, graphs, ngraphs, queries, nqueries, max_workgroups));
  return batch_wait(ctx);
}

}  // namespace indigo
