// Batched engine: many graphs and many sources per graph in one persistent
// launch (MI355X, gfx950).
//
// Every query becomes a task: multi-source BFS (up to 64 sources), multi-
// source SSSP (up to 32), CC, MIS or PageRank on one graph.  The host sorts
// tasks by size (largest first) into a device-side queue; one persistent
// kernel, sized to co-residency and checked by a residency barrier, serves
// it.
//
// Gangs.  A task runs on a gang of k workgroups, k = the task's CSR entries
// / 2^18 rounded up to a power of two: small graphs take one workgroup and
// synchronise with __syncthreads only, large ones a gang with its own bounded
// barrier.  The cap depends on how much work a barrier separates: 256 for
// the dense sweeps of PR, CC and MIS, 128 for BFS, 64 for SSSP's many short
// rounds, and 16 for BFS on deep sparse graphs (average degree below 6),
// whose thousands of levels are bound by barrier latency.  A free workgroup joins the first task in queue
// order that still has a free slot (one atomic on the task's `joined`
// counter); the gang starts when all k members have arrived.  Because every
// workgroup is resident and gangs are filled strictly in queue order, a gang
// only ever waits for workgroups that are busy with earlier tasks, so
// formation cannot deadlock.
//
// Load balance inside a task.  Sweeps give each wave 64 consecutive vertices
// (or queue entries) and walk their lists 64 entries per step with the
// in-wave balancer.  Lists of kHeavyDeg entries or more would tie a wave up
// (R-MAT hubs reach 10^5 entries), so each task first lists its heavy
// vertices as kItem-entry items that all waves of the gang share; sweeps
// skip heavy lists and a second loop walks the items.
//
// Memory discipline (as the coherent-mode persistent kernels): data one
// thread writes and another reads within a task, or that two threads write,
// goes through device-scope relaxed loads / stores / atomics (the eight XCDs'
// L2 caches are not coherent for plain accesses, and these barriers do not
// write them back), and each barrier first drains the wave's memory
// operations; barriers are relaxed counters.  Plain stores remain only for
// slab entries written exactly once.
//
// Algorithms (same results as the single-query kernels and the reference):
//   BFS : level-synchronous and direction-optimising over 64-bit masks:
//         frontier[v] holds the sources whose BFS reached v at the previous
//         level, and the frontier vertices are also kept in a queue.  While
//         the frontier's lists hold fewer than E / 14 entries it pushes:
//         each frontier vertex ORs its mask, minus the bits a neighbour has
//         seen, into the neighbour's next mask, and the first push into a
//         vertex queues it.  Otherwise it pulls: every vertex with sources
//         still missing ORs its neighbours' frontier masks.  A second phase
//         turns the next masks into the new frontier and writes the level
//         into the slab rows of the new bits.  Deep graphs (grids, roads)
//         cost their edge work plus two barriers per level.
//   SSSP: Bellman-Ford driven by a queue of active vertices over a
//         per-source distance slab, with per-vertex 32-bit active masks;
//         relaxations by atomicMin, a vertex is queued for the next round
//         when its distance for some source drops.
//   CC  : union-find hooking the larger root under the smaller (labels =
//         component minima), then flattening; graphs of average degree 8 or
//         more link in Afforest order and skip the giant component's lists.
//   MIS : the greedy fixed point in descending (hash(v + 712313887), v)
//         order; per round a decide phase (a vertex joins when no neighbour
//         is included and no undecided neighbour has a larger key) and an
//         exclude phase.  A heavy vertex's items raise per-vertex flags
//         (blocked, has an included neighbour) in the decide phase, and the
//         vertex is decided from them in the exclude phase.  Deferring a
//         decision by a phase changes no result: a vertex decides only once
//         every larger undecided neighbour has, as in the greedy order.
//   PR  : the Jacobi recurrence of the single-query kernel (FP32 scores,
//         Neumaier-compensated gathers: one lane per list of fewer than
//         kPrWaveList entries in entry order, one wave per longer one,
//         per-item partials combined in item order for heavy lists),
//         stopping when the FP64 L1 change is <= 1e-4 or after 100 sweeps;
//         every gang member adds the members' FP64 partials in the same
//         order, so all take the same decision.
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
constexpr int kMaxGang = INDIGO_MAX_GANG;
constexpr int kGangShift = 18;   // CSR entries per gang member
constexpr int kHeavyDeg = 1024;  // lists this long are split into items
constexpr int kItem = 1024;      // entries per item
constexpr int kCcSample = 2;      // CC: neighbours linked per vertex before the giant root is chosen
constexpr int kCcSamples = 256;   // CC: vertices sampled to choose it
constexpr int kCcAfforestMin = 4096;  // CC: smallest graph that takes the Afforest order
constexpr int kPrWaveList = 32;  // PR: lists walked by a whole wave
constexpr float kDamp = 0.85f;

struct alignas(128) GangState {
  unsigned long long mf[4];  // BFS: degree sums of the queued frontiers, rotated by level
  unsigned joined;           // members that took a slot (grows past the gang size)
  unsigned arrive;           // barrier arrivals of the current generation
  unsigned gen;              // barrier generation
  unsigned flag[3];          // per-round "something changed" flags, triple-buffered
  unsigned qn[4];            // frontier queue sizes, rotated by level / round
  unsigned nitems;           // heavy-list items
  unsigned grp[8];           // barrier arrivals per rank group (gangs above 16)
  unsigned pad[5];
};

struct BatchTask {
  int alg, gang, nodes, edges;
  const int* nindex;
  const int* nlist;
  const int* eweight;
  int* out;
  void* scratch;  // per algorithm, see scratch_bytes
  int2* items;    // heavy-list items (vertex, first entry)
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

// Gang barrier (bounded).  Size 1: the workgroup barrier suffices; up to 64
// members count on one counter, larger gangs first on a counter per rank
// group (rank mod 8), then the groups' last arrivals on the shared one.
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
      bool last = false;
      if (tm.size <= 64) {
        last = atomic_add_relaxed(&tm.gs->arrive, 1u) == static_cast<unsigned>(tm.size) - 1u;
      } else {
        const unsigned g = static_cast<unsigned>(tm.rank) & 7u;
        const unsigned gsize = (static_cast<unsigned>(tm.size) - g + 7u) >> 3;  // ranks r < size with r % 8 == g
        if (atomic_add_relaxed(&tm.gs->grp[g], 1u) == gsize - 1u) {
          st_relaxed(&tm.gs->grp[g], 0u);
          wave_drain();
          last = atomic_add_relaxed(&tm.gs->arrive, 1u) == 7u;
        }
      }
      if (last) {
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
          for (int k = 0; k < backoff; k++) __builtin_amdgcn_s_sleep(1);
          backoff = backoff < 8 ? backoff * 2 : 8;  // short polls: gangs wait a few microseconds per level
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

IG_DEV bool heavy(int deg) { return deg >= kHeavyDeg; }

// Wave-collective: walks the CSR entries of the vertex `v` of each lane with
// `active` and a list shorter than kHeavyDeg, 64 per step; f(o, e, valid)
// gets the owner lane o and the entry e, with all lanes converged.
template <typename F>
IG_DEV void wave_entries_of(int v, bool active, const int* __restrict__ nindex, F&& f)
{
  const int lane = lane_id();
  const int b = active ? nindex[v] : 0;
  const int l = active ? nindex[v + 1] - b : 0;
  const int len = heavy(l) ? 0 : l;
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

IG_DEV Chunks team_chunks(const Team& tm, int n)
{
  return Chunks{tm.rank * kBatchWaves + static_cast<int>(threadIdx.x) / kWave, tm.size * kBatchWaves,
                (n + kWave - 1) / kWave};
}

// Wave-collective: f(v, e_begin, e_end, item) for the team's share of the
// heavy items (one item per wave at a time).
template <typename F>
IG_DEV void for_items(const BatchTask& T, const Team& tm, int nitems, F&& f)
{
  for (int i = tm.rank * kBatchWaves + static_cast<int>(threadIdx.x) / kWave; i < nitems;
       i += tm.size * kBatchWaves) {
    const int v = ld_relaxed(&T.items[i].x), eb = ld_relaxed(&T.items[i].y);
    f(v, eb, min(eb + kItem, T.nindex[v + 1]), i);
  }
}

// Gang-collective, once per task: lists the heavy vertices' items.
IG_DEV bool build_items(const BatchTask& T, const Team& tm, int* nitems)
{
  const int lane = lane_id();
  const Chunks ch = team_chunks(tm, T.nodes);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    const int b = v < T.nodes ? T.nindex[v] : 0;
    const int deg = v < T.nodes ? T.nindex[v + 1] - b : 0;
    const int n = heavy(deg) ? (deg + kItem - 1) / kItem : 0;
    const int incl = wave_inclusive_scan(n);
    const int total = wave_bcast(incl, kWave - 1);
    if (total == 0) continue;
    unsigned base = 0;
    if (lane == 0) base = atomic_add_relaxed(&tm.gs->nitems, static_cast<unsigned>(total));
    base = static_cast<unsigned>(wave_shfl(static_cast<int>(base), 0));
    for (int j = 0; j < n; j++) {
      const unsigned slot = base + static_cast<unsigned>(incl - n + j);
      st_relaxed(&T.items[slot].x, v);
      st_relaxed(&T.items[slot].y, b + j * kItem);
    }
  }
  if (!team_sync(tm)) return false;
  *nitems = static_cast<int>(ld_relaxed(&tm.gs->nitems));
  return true;
}

IG_DEV void set_flag(unsigned* flag, bool any)
{
  if (__ballot(any) != 0ull && lane_id() == 0) st_relaxed(flag, 1u);
}

IG_DEV unsigned long long wave_or(unsigned long long x)
{
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x |= __shfl_xor(x, d, kWave);
  return x;
}

// Wave-collective append of u (where pred) to a queue, adding the degrees of
// the appended vertices to *mf.
IG_DEV void append_frontier(bool pred, int u, int deg, int* queue, unsigned* count, unsigned long long* mf)
{
  wave_append(pred, u, queue, count);
  const unsigned long long d = wave_sum(static_cast<unsigned long long>(pred ? deg : 0));
  if (lane_id() == 0 && d != 0ull) atomic_add_relaxed(mf, d);
}

// ---- multi-source BFS -------------------------------------------------------
IG_DEV bool run_bfs(const BatchTask& T, const Team& tm, unsigned long long (*s_acc)[kWave])
{
  const int V = T.nodes;
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const size_t Vs = static_cast<size_t>(V);
  unsigned long long* seen = static_cast<unsigned long long*>(T.scratch);
  unsigned long long* fr[2] = {seen + Vs, seen + 2 * Vs};  // frontier masks by level parity
  unsigned long long* nxt = seen + 3 * Vs;                 // next-level masks
  int* queue[2] = {reinterpret_cast<int*>(seen + 4 * Vs), reinterpret_cast<int*>(seen + 4 * Vs) + V};
  unsigned* qn = tm.gs->qn;
  unsigned long long* mf = tm.gs->mf;
  const unsigned long long all = T.nsrc >= 64 ? ~0ull : ((1ull << T.nsrc) - 1ull);
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) {
      st_relaxed(&seen[v], 0ull);
      st_relaxed(&fr[0][v], 0ull);
      st_relaxed(&fr[1][v], 0ull);
      st_relaxed(&nxt[v], 0ull);
    }
  }
  int nitems = 0;
  if (!build_items(T, tm, &nitems)) return false;
  if (tm.rank == 0 && threadIdx.x < kWave) {  // one wave queues the distinct sources
    const bool has = lane < T.nsrc;
    const int i = lane;
    const int s = has ? T.src[i] : 0;
    unsigned long long old = 0ull;
    if (has) {
      __hip_atomic_fetch_or(&seen[s], 1ull << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      old = __hip_atomic_fetch_or(&fr[0][s], 1ull << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st_relaxed(&T.out[static_cast<size_t>(i) * V + s], 0);
    }
    // a source listed twice is queued by the lane that set its first bit
    append_frontier(has && old == 0ull, s, has ? T.nindex[s + 1] - T.nindex[s] : 0, queue[0], &qn[0], &mf[0]);
  }
  if (!team_sync(tm)) return false;
  for (int lvl = 1;; lvl++) {
    const unsigned long long* cur = fr[(lvl - 1) & 1];
    unsigned long long* nfr = fr[lvl & 1];
    const int* qcur = queue[(lvl - 1) & 1];
    int* qnext = queue[lvl & 1];
    unsigned* ncount = &qn[lvl & 3];
    unsigned long long* nmf = &mf[lvl & 3];
    const int ncur = static_cast<int>(ld_relaxed(&qn[(lvl - 1) & 3]));
    const unsigned long long mcur = ld_relaxed(&mf[(lvl - 1) & 3]);
    if (tm.rank == 0 && threadIdx.x == 0) {
      st_relaxed(&qn[(lvl + 1) & 3], 0u);
      st_relaxed(&mf[(lvl + 1) & 3], 0ull);
    }
    if (mcur * 14ull < static_cast<unsigned long long>(T.edges)) {
      // push from the frontier queue
      const int nq = (ncur + kWave - 1) / kWave;
      auto push = [&](unsigned long long m, int e, bool valid) {
        bool first = false;
        int u = 0;
        if (valid) {
          u = T.nlist[e];
          const unsigned long long add = m & ~ld_relaxed(&seen[u]);
          if (add != 0ull)
            first = __hip_atomic_fetch_or(&nxt[u], add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull;
        }
        append_frontier(first, u, first ? T.nindex[u + 1] - T.nindex[u] : 0, qnext, ncount, nmf);
      };
      for (int c = ch.first; c < nq; c += ch.step) {
        const int k = c * kWave + lane;
        const bool act = k < ncur;
        const int v = act ? ld_relaxed(&qcur[k]) : 0;
        const unsigned long long m = act ? ld_relaxed(&cur[v]) : 0ull;
        wave_entries_of(v, act, T.nindex, [&](int o, int e, bool valid) { push(__shfl(m, o, kWave), e, valid); });
      }
      for_items(T, tm, nitems, [&](int v, int eb, int ee, int) {
        const unsigned long long m = ld_relaxed(&cur[v]);
        if (m == 0ull) return;
        for (int e0 = eb; e0 < ee; e0 += kWave) push(m, e0 + lane, e0 + lane < ee);
      });
    } else {
      // pull: every vertex with sources still missing
      for (int c = ch.first; c < ch.count; c += ch.step) {
        const int v = c * kWave + lane;
        const unsigned long long sv = v < V ? ld_relaxed(&seen[v]) : all;
        const unsigned long long missing = all & ~sv;
        s_acc[wid][lane] = 0ull;
        wave_lds_sync();
        wave_entries_of(v, missing != 0ull, T.nindex, [&](int o, int e, bool valid) {
          if (valid) {
            const unsigned long long m = ld_relaxed(&cur[T.nlist[e]]);
            if (m != 0ull) atomicOr(&s_acc[wid][o], m);
          }
        });
        wave_lds_sync();
        const unsigned long long nb = s_acc[wid][lane] & missing;
        if (nb != 0ull) st_relaxed(&nxt[v], nb);
        append_frontier(nb != 0ull, v, nb != 0ull ? T.nindex[v + 1] - T.nindex[v] : 0, qnext, ncount, nmf);
      }
      for_items(T, tm, nitems, [&](int v, int eb, int ee, int) {
        const unsigned long long missing = all & ~ld_relaxed(&seen[v]);
        if (missing == 0ull) return;
        unsigned long long acc = 0ull;
        for (int e = eb + lane; e < ee; e += kWave) acc |= ld_relaxed(&cur[T.nlist[e]]);
        const unsigned long long nb = wave_or(acc) & missing;
        bool first = false;
        if (nb != 0ull && lane == 0)
          first = __hip_atomic_fetch_or(&nxt[v], nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull;
        append_frontier(first, v, first ? T.nindex[v + 1] - T.nindex[v] : 0, qnext, ncount, nmf);
      });
    }
    if (!team_sync(tm)) return false;
    // the new frontier: masks, seen bits and levels; the old frontier's masks
    // are cleared (that buffer holds the next level's frontier)
    const int nnext = static_cast<int>(ld_relaxed(ncount));
    for (int k = tm.rank * kBatchBlock + static_cast<int>(threadIdx.x); k < ncur; k += tm.size * kBatchBlock)
      st_relaxed(&fr[(lvl - 1) & 1][ld_relaxed(&qcur[k])], 0ull);
    for (int k = tm.rank * kBatchBlock + static_cast<int>(threadIdx.x); k < nnext; k += tm.size * kBatchBlock) {
      const int u = ld_relaxed(&qnext[k]);
      const unsigned long long sv = ld_relaxed(&seen[u]);
      const unsigned long long nb = ld_relaxed(&nxt[u]) & ~sv;
      st_relaxed(&nxt[u], 0ull);
      st_relaxed(&seen[u], sv | nb);
      st_relaxed(&nfr[u], nb);
      for (unsigned long long m = nb; m != 0ull; m &= m - 1ull)
        T.out[static_cast<size_t>(__ffsll(static_cast<long long>(m)) - 1) * V + u] = lvl;
    }
    if (!team_sync(tm)) return false;
    if (nnext == 0) break;
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
  const size_t Vs = static_cast<size_t>(V);
  unsigned* base = static_cast<unsigned*>(T.scratch);
  unsigned* act[3] = {base, base + Vs, base + 2 * Vs};  // active masks by round mod 3
  int* queue[3] = {reinterpret_cast<int*>(base + 3 * Vs), reinterpret_cast<int*>(base + 4 * Vs),
                   reinterpret_cast<int*>(base + 5 * Vs)};
  unsigned* qn = tm.gs->qn;
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v >= V) continue;
    st_relaxed(&act[0][v], 0u);
    st_relaxed(&act[1][v], 0u);
    st_relaxed(&act[2][v], 0u);
    for (int i = 0; i < T.nsrc; i++) st_relaxed(&T.out[static_cast<size_t>(i) * V + v], INT_MAX);
  }
  int nitems = 0;
  if (!build_items(T, tm, &nitems)) return false;
  if (tm.rank == 0 && threadIdx.x < kWave) {
    const bool has = lane < T.nsrc;
    const int i = lane;
    const int s = has ? T.src[i] : 0;
    unsigned old = 0u;
    if (has) {
      st_relaxed(&T.out[static_cast<size_t>(i) * V + s], 0);
      old = __hip_atomic_fetch_or(&act[0][s], 1u << i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    wave_append(has && old == 0u, s, queue[0], &qn[0]);
  }
  if (!team_sync(tm)) return false;
  for (int r = 0;; r++) {
    const unsigned* cur = act[r % 3];
    unsigned* nxt = act[(r + 1) % 3];
    unsigned* prev = act[(r + 2) % 3];  // round r-1's masks: cleared now
    const int* qcur = queue[r % 3];
    int* qnext = queue[(r + 1) % 3];
    const int* qprev = queue[(r + 2) % 3];
    unsigned* ncount = &qn[(r + 1) & 3];
    const int ncur = static_cast<int>(ld_relaxed(&qn[r & 3]));
    const int nprev = r > 0 ? static_cast<int>(ld_relaxed(&qn[(r + 3) & 3])) : 0;
    if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&qn[(r + 2) & 3], 0u);
    for (int k = tm.rank * kBatchBlock + static_cast<int>(threadIdx.x); k < nprev; k += tm.size * kBatchBlock)
      st_relaxed(&prev[ld_relaxed(&qprev[k])], 0u);
    auto relax = [&](unsigned m, int v, int e, bool valid) {
      bool first = false;
      int u = 0;
      if (valid) {
        u = T.nlist[e];
        const int w = T.eweight[e];
        unsigned gained = 0u;
        for (unsigned bits = m; bits != 0u; bits &= bits - 1u) {
          const int i = __ffs(static_cast<int>(bits)) - 1;
          const size_t row = static_cast<size_t>(i) * V;
          const int nd = ld_relaxed(&T.out[row + v]) + w;
          if (nd < ld_relaxed(&T.out[row + u]) && nd < atomic_min_relaxed(&T.out[row + u], nd)) gained |= 1u << i;
        }
        if (gained != 0u)
          first = __hip_atomic_fetch_or(&nxt[u], gained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
      }
      wave_append(first, u, qnext, ncount);
    };
    const int nq = (ncur + kWave - 1) / kWave;
    for (int c = ch.first; c < nq; c += ch.step) {
      const int k = c * kWave + lane;
      const bool a = k < ncur;
      const int v = a ? ld_relaxed(&qcur[k]) : 0;
      const unsigned m = a ? ld_relaxed(&cur[v]) : 0u;
      wave_entries_of(v, a, T.nindex, [&](int o, int e, bool valid) {
        relax(static_cast<unsigned>(__shfl(static_cast<int>(m), o, kWave)), __shfl(v, o, kWave), e, valid);
      });
    }
    for_items(T, tm, nitems, [&](int v, int eb, int ee, int) {
      const unsigned m = ld_relaxed(&cur[v]);
      if (m == 0u) return;
      for (int e0 = eb; e0 < ee; e0 += kWave) relax(m, v, e0 + lane, e0 + lane < ee);
    });
    if (!team_sync(tm)) return false;
    if (ld_relaxed(ncount) == 0u) break;
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

// Root of x without path halving.  The final flatten writes roots into the
// same array the finds walk: a halving store that read parent[x] before it
// was flattened could put an intermediate ancestor back over the root, so
// that pass stores roots only.
IG_DEV int root_of(const int* parent, int x)
{
  int p;
  while ((p = ld_relaxed(&parent[x])) != x) x = p;
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

// Workgroup-collective: the most frequent root among kCcSamples vertices
// spread over the id range (ties to the smaller root).  Every gang member
// reads the same compressed parents, so all choose the same root.
IG_DEV int giant_root(const BatchTask& T, int* parent, unsigned long long (*s_acc)[kWave])
{
  int* s_root = reinterpret_cast<int*>(&s_acc[0][0]);  // kCcSamples ints of the task scratch LDS
  const int V = T.nodes;
  for (int i = threadIdx.x; i < kCcSamples; i += blockDim.x)
    s_root[i] = find_root(parent, static_cast<int>((static_cast<unsigned long long>(i) * 2654435761ull) %
                                                   static_cast<unsigned>(V)));
  __syncthreads();
  unsigned long long key = 0;  // (count << 32) | ~root: the largest wins
  for (int i = threadIdx.x; i < kCcSamples; i += blockDim.x) {
    unsigned c = 0;
    for (int j = 0; j < kCcSamples; j++) c += s_root[j] == s_root[i] ? 1u : 0u;
    const unsigned long long k = (static_cast<unsigned long long>(c) << 32) | (~static_cast<unsigned>(s_root[i]));
    key = k > key ? k : key;
  }
  __syncthreads();
  key = wave_max(key);
  if (lane_id() == 0) s_acc[threadIdx.x / kWave][0] = key;
  __syncthreads();
  unsigned long long best = 0;
  for (int w = 0; w < kBatchWaves; w++) best = s_acc[w][0] > best ? s_acc[w][0] : best;
  __syncthreads();
  return static_cast<int>(~static_cast<unsigned>(best & 0xffffffffull));
}

// CC as the single-query kernel: union-find hooking the larger root under
// the smaller (labels = component minima).  Graphs with an average degree of
// at least 8 (and kCcAfforestMin vertices) take the Afforest order - link
// each vertex with its first kCcSample neighbours, compress, choose the giant
// root, and let only vertices outside its tree link their lists (all
// entries); the others link every undirected edge once, from its larger
// endpoint.
IG_DEV bool run_cc(const BatchTask& T, const Team& tm, unsigned long long (*s_acc)[kWave])
{
  const int V = T.nodes;
  const int lane = lane_id();
  int* parent = T.out;
  const bool aff = static_cast<long long>(T.edges) >= 8LL * V && V >= kCcAfforestMin;
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) {
      int p = v;
      if (aff && T.nindex[v] < T.nindex[v + 1]) p = min(v, T.nlist[T.nindex[v]]);
      st_relaxed(&parent[v], p);
    }
  }
  int nitems = 0;
  if (!build_items(T, tm, &nitems)) return false;  // (ends with a team barrier)
  int giant = -1;
  if (aff) {
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      if (v < V) {
        const int b = T.nindex[v], e = min(T.nindex[v + 1], b + kCcSample);
        for (int k = b; k < e; k++) link(parent, v, T.nlist[k]);
      }
    }
    if (!team_sync(tm)) return false;
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      if (v < V) st_relaxed(&parent[v], find_root(parent, v));
    }
    if (!team_sync(tm)) return false;
    giant = giant_root(T, parent, s_acc);
  }
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    bool todo = v < V;
    if (aff && todo) todo = T.nindex[v + 1] - T.nindex[v] > kCcSample && find_root(parent, v) != giant;
    wave_entries_of(v, todo, T.nindex, [&](int o, int e, bool valid) {
      if (!valid) return;
      const int ov = c * kWave + o;
      const int u = T.nlist[e];
      if (aff || u < ov) link(parent, ov, u);
    });
  }
  for_items(T, tm, nitems, [&](int v, int eb, int ee, int) {
    if (aff && find_root(parent, v) == giant) return;
    for (int e = eb + lane; e < ee; e += kWave) {
      const int u = T.nlist[e];
      if (aff || u < v) link(parent, v, u);
    }
  });
  if (!team_sync(tm)) return false;
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) st_relaxed(&parent[v], root_of(parent, v));
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
  unsigned* blocked = static_cast<unsigned*>(T.scratch);  // heavy vertices: a blocking neighbour seen
  unsigned* near_in = blocked + V;                         // heavy vertices: an included neighbour seen
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v < V) {
      st_relaxed(&status[v], 0);
      st_relaxed(&blocked[v], 0u);
      st_relaxed(&near_in[v], 0u);
    }
  }
  int nitems = 0;
  if (!build_items(T, tm, &nitems)) return false;
  for (int r = 0;; r++) {
    unsigned* flag = &tm.gs->flag[r % 3];
    if (tm.rank == 0 && threadIdx.x == 0) st_relaxed(&tm.gs->flag[(r + 1) % 3], 0u);
    // decide (lists shorter than kHeavyDeg) and gather the heavy vertices' flags
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const int deg = v < V ? T.nindex[v + 1] - T.nindex[v] : 0;
      const bool und = v < V && !heavy(deg) && ld_relaxed(&status[v]) == 0;
      s_flag[wid][lane] = 0;
      wave_lds_sync();
      wave_entries_of(v, und, T.nindex, [&](int o, int e, bool valid) {
        if (!valid) return;
        const int ov = c * kWave + o;
        const int u = T.nlist[e];
        // only larger keys can block (or be included next to) an undecided
        // vertex: a smaller one cannot be included before it is decided
        if (u != ov && mis_key(u) > mis_key(ov) && ld_relaxed(&status[u]) != 2) s_flag[wid][o] = 1;
      });
      wave_lds_sync();
      if (und && s_flag[wid][lane] == 0) st_relaxed(&status[v], 1);
    }
    for_items(T, tm, nitems, [&](int v, int eb, int ee, int) {
      if (ld_relaxed(&status[v]) != 0) return;
      const unsigned long long kv = mis_key(v);
      bool blk = false, inc = false;
      for (int e = eb + lane; e < ee; e += kWave) {
        const int u = T.nlist[e];
        if (u == v || mis_key(u) < kv) continue;  // smaller keys cannot decide v
        const int su = ld_relaxed(&status[u]);
        inc |= su == 1;
        blk |= su != 2;
      }
      const bool any_blk = __ballot(blk) != 0ull, any_inc = __ballot(inc) != 0ull;
      if (lane == 0 && any_blk) st_relaxed(&blocked[v], 1u);
      if (lane == 0 && any_inc) st_relaxed(&near_in[v], 1u);
    });
    if (!team_sync(tm)) return false;
    // exclude the undecided neighbours of included vertices; decide the heavy
    // vertices from their flags
    for (int c = ch.first; c < ch.count; c += ch.step) {
      const int v = c * kWave + lane;
      const int deg = v < V ? T.nindex[v + 1] - T.nindex[v] : 0;
      const bool und = v < V && ld_relaxed(&status[v]) == 0;
      const bool hv = und && heavy(deg);
      s_flag[wid][lane] = 0;
      wave_lds_sync();
      wave_entries_of(v, und && !hv, T.nindex, [&](int o, int e, bool valid) {
        if (!valid) return;
        const int u = T.nlist[e];
        // an included neighbour of an undecided vertex has a larger key
        if (mis_key(u) > mis_key(c * kWave + o) && ld_relaxed(&status[u]) == 1) s_flag[wid][o] = 1;
      });
      wave_lds_sync();
      bool still = false;
      if (und && !hv) {
        if (s_flag[wid][lane] != 0) st_relaxed(&status[v], 2);
        else still = true;
      }
      if (hv) {
        if (ld_relaxed(&near_in[v]) != 0u) st_relaxed(&status[v], 2);
        else if (ld_relaxed(&blocked[v]) == 0u) st_relaxed(&status[v], 1);
        else still = true;
        st_relaxed(&blocked[v], 0u);
        st_relaxed(&near_in[v], 0u);
      }
      set_flag(flag, still);
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
  const size_t Vs = static_cast<size_t>(V);
  float* scores = reinterpret_cast<float*>(T.out);
  double* slots = static_cast<double*>(T.scratch);  // [2][kMaxGang] FP64 partials
  float* contrib[2] = {reinterpret_cast<float*>(slots + 2 * kMaxGang),
                       reinterpret_cast<float*>(slots + 2 * kMaxGang) + Vs};
  float2* hpart = reinterpret_cast<float2*>(contrib[1] + Vs);  // per item (sum, error)
  const float inv_n = 1.0f / static_cast<float>(V);
  const float base = (1.0f - kDamp) / static_cast<float>(V);
  const Chunks ch = team_chunks(tm, V);
  for (int c = ch.first; c < ch.count; c += ch.step) {
    const int v = c * kWave + lane;
    if (v >= V) continue;
    st_relaxed(&scores[v], inv_n);
    st_relaxed(&contrib[0][v], inv_n / static_cast<float>(T.nindex[v + 1] - T.nindex[v]));
  }
  int nitems = 0;
  if (!build_items(T, tm, &nitems)) return false;
  // a heavy vertex's score is written at init by its chunk's wave and later by
  // the wave of its first item, possibly on another XCD: device-scope stores
  // keep the two L2s from writing back stale bytes
  auto finish = [&](float* cout, int x, int xdeg, float total) {
    const float old = ld_relaxed(&scores[x]);
    const float s = base + kDamp * total;
    st_relaxed(&scores[x], s);
    st_relaxed(&cout[x], s / static_cast<float>(xdeg));
    return fabs(static_cast<double>(s) - static_cast<double>(old));
  };
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
      if (v < V && deg < kPrWaveList) {
        CSum acc;
        for (int e = beg; e < beg + deg; e++) acc.add(ld_relaxed(&cin[T.nlist[e]]));
        d += finish(cout, v, deg, acc.value());
      }
      for (unsigned long long m = __ballot(v < V && deg >= kPrWaveList && !heavy(deg)); m != 0ull; m &= m - 1ull) {
        const int o = __ffsll(static_cast<long long>(m)) - 1;
        const int x = c * kWave + o;
        const int xb = wave_bcast(beg, o), xd = wave_bcast(deg, o);
        CSum acc;
        for (int e = xb + lane; e < xb + xd; e += kWave) acc.add(ld_relaxed(&cin[T.nlist[e]]));
        acc = wave_merge(acc);
        if (lane == 0) d += finish(cout, x, xd, acc.value());
      }
    }
    if (nitems > 0) {
      for_items(T, tm, nitems, [&](int, int eb, int ee, int i) {
        CSum acc;
        for (int e = eb + lane; e < ee; e += kWave) acc.add(ld_relaxed(&cin[T.nlist[e]]));
        acc = wave_merge(acc);
        if (lane == 0) {
          st_relaxed(&hpart[i].x, acc.s);
          st_relaxed(&hpart[i].y, acc.c);
        }
      });
      if (!team_sync(tm)) return false;
      // one wave per heavy vertex (at its first item) combines its items in order
      for_items(T, tm, nitems, [&](int v, int eb, int, int i) {
        if (eb != T.nindex[v]) return;
        const int deg = T.nindex[v + 1] - eb;
        const int n = (deg + kItem - 1) / kItem;
        CSum acc;
        for (int j = lane; j < n; j += kWave) acc.merge(ld_relaxed(&hpart[i + j].x), ld_relaxed(&hpart[i + j].y));
        acc = wave_merge(acc);
        if (lane == 0) d += finish(cout, v, deg, acc.value());
      });
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

__global__ __launch_bounds__(kBatchBlock) IG_SGPR_CAP void batch_kernel(const BatchTask* __restrict__ tasks, int ntasks,
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
    else if (T.alg == INDIGO_QUERY_CC) ok = run_cc(T, tm, s_acc);
    else if (T.alg == INDIGO_QUERY_PR) ok = run_pr(T, tm, s_red);
    else ok = run_mis(T, tm, s_flag);
    if (!ok) break;
  }
}

size_t items_cap(int edges) { return 2 * (static_cast<size_t>(edges) / kItem) + 2; }

// Per-task scratch: the algorithm's arrays (alg_bytes), then the heavy items.
size_t alg_bytes(int alg, int nodes, int edges)
{
  const size_t v = static_cast<size_t>(nodes);
  size_t b = 0;
  if (alg == INDIGO_QUERY_BFS) b = 4 * v * 8 + 2 * v * 4;             // seen, 2 frontiers, next; 2 queues
  else if (alg == INDIGO_QUERY_SSSP) b = 6 * v * 4;                  // 3 active masks, 3 queues
  else if (alg == INDIGO_QUERY_MIS) b = 2 * v * 4;                   // heavy-vertex flags
  else if (alg == INDIGO_QUERY_PR) b = 2 * kMaxGang * 8 + 2 * v * 4 + items_cap(edges) * 8;  // slots, contrib, partials
  return Carver::round(b);
}

size_t scratch_bytes(int alg, int nodes, int edges)
{
  return alg_bytes(alg, nodes, edges) + Carver::round(items_cap(edges) * sizeof(int2));
}

}  // namespace

int batch_gang(int alg, int nodes, int edges)
{
  if (alg < INDIGO_QUERY_BFS || alg > INDIGO_QUERY_PR || nodes < 0 || edges < 0) return 0;
  const long long entries = static_cast<long long>(edges) + nodes;
  int cap = kMaxGang;
  if (alg == INDIGO_QUERY_BFS || alg == INDIGO_QUERY_SSSP) {
    cap = alg == INDIGO_QUERY_BFS ? 128 : 64;
    if (alg == INDIGO_QUERY_BFS && static_cast<long long>(edges) < 6LL * nodes) cap = 16;
  }
  int k = 1;
  while (k < cap && (static_cast<long long>(k) << kGangShift) < entries) k <<= 1;
  return k;
}

namespace {

int gang_for(int alg, int nodes, int edges, int grid) { return std::min(batch_gang(alg, nodes, edges), std::max(1, grid)); }

}  // namespace

// Workgroups that every persistent kernel of the library can hold resident
// at once: the batch kernel's occupancy with one workgroup per CU left out,
// as for the kernels of more than 80 scalar registers (BFS, SSSP, PageRank
// stop at 3 x 512 threads per CU), so that shares of it assemble side by
// side whatever kernels the sharers run.
int batch_grid_limit(indigo_ctx* ctx, int* limit)
{
  int per_cu = 0;
  IG_TRY(resident_per_cu(ctx, reinterpret_cast<const void*>(batch_kernel), kBatchBlock, 0, &per_cu, false));
  *limit = std::max(1, per_cu * ctx->compute_units);
  return INDIGO_OK;
}

// (Re)launches the batch recorded in ctx->pending_batch; records the stop
// event without waiting.
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
                     ctx->pending_batch.grid_cap, nullptr, "batch_kernel", true));
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
    if (Q.gang < 0 || Q.gang > kMaxGang)
      return set_error(INDIGO_ERR_INVALID_ARG, "batch: query %d gang %d outside 0..%d", q, Q.gang, kMaxGang);
  }
  // grid: occupancy x CUs (capped), learned residency caps as for the other persistent kernels
  int per_cu = 0;
  IG_TRY(resident_per_cu(ctx, reinterpret_cast<const void*>(batch_kernel), kBatchBlock, 0, &per_cu, true));
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
  for (int q = 0; q < nqueries; q++)
    scratch_total += scratch_bytes(queries[q].alg, graphs[queries[q].graph].nodes, graphs[queries[q].graph].edges);
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
    T.gang = Q.gang > 0 ? std::min(Q.gang, grid_cap) : gang_for(Q.alg, G.nodes, G.edges, grid_cap);
    T.nodes = G.nodes;
    T.edges = G.edges;
    T.nindex = G.d_nindex;
    T.nlist = G.d_nlist;
    T.eweight = G.d_eweight;
    T.out = Q.d_out;
    T.scratch = scratch + off;
    T.items = reinterpret_cast<int2*>(scratch + off + alg_bytes(Q.alg, G.nodes, G.edges));
    off += scratch_bytes(Q.alg, G.nodes, G.edges);
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
