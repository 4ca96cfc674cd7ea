// Degree-binned frontier queues for the persistent traversal kernels.
//
// Queue contents and counters are written and read with device-scope atomic
// accesses only, so kernels built on them may use the coherent grid barrier.
//
// A frontier is three queues, binned by adjacency-list length when a vertex
// is appended:
//   * small: length < kMidDeg.  Expanded 64 vertices per wave with an in-wave
//     load balancer (prefix sum of lengths + owner search), so a wave always
//     handles 64 consecutive CSR entries and never more than 64 * 63 in all;
//   * mid:   kMidDeg <= length < kHugeDeg.  One wave per vertex, lanes
//     striding its list (coalesced);
//   * huge:  length >= kHugeDeg.  Split into kHugeChunk-entry work items
//     spread over all workgroups.  Its counter packs (chunk count << 32 |
//     vertex count) so one 64-bit atomic allocates both the slot and the
//     vertex's first chunk index; hoff[] is therefore sorted by slot and a
//     chunk's owner is found by binary search.
// Appends are wave-aggregated (ballot + mbcnt prefix, one atomic per wave and
// queue).  Counters are triple-buffered across levels so that the set written
// in level L, read in L+1 and cleared in L+2 never has two roles at once.
//
// The same machinery bins a static vertex set: collect_heavy() appends every
// vertex of length >= kMidDeg of a graph to the mid / huge queues, while the
// short lists are walked in vertex-id order (coalesced) by the kernels.
#pragma once

#include "device_common.hpp"
#include "emit.hpp"

namespace indigo {

constexpr int kMidDeg = 64;
constexpr int kHugeDeg = 4096;
constexpr int kHugeChunk = 2048;

struct alignas(128) FrontierCounters {
  unsigned int n;            // small queue length
  unsigned int nm;           // mid queue length
  unsigned long long hpack;  // (huge chunks << 32) | huge queue length
  unsigned long long m;      // sum of lengths over the frontier
  unsigned long long aux;    // algorithm-specific
  unsigned long long pad[12];
};

// Per-level trace written by workgroup 0 of a traversal kernel: wall-clock
// tick at the start of the level, frontier size, its degree sum and the mode
// (0 top-down, 1 bottom-up, 2 one-workgroup phase, 3 far-pile split).
constexpr int kTraceMax = 256;
struct LevelTrace {
  int n;
  int pad;
  long long start;
  long long t[kTraceMax];
  unsigned nf[kTraceMax];
  unsigned long long m[kTraceMax];
  int mode[kTraceMax];
};

IG_DEV void trace_level(LevelTrace* tr, int level, unsigned nf, unsigned long long m, int mode)
{
  if (blockIdx.x == 0 && threadIdx.x == 0 && level < kTraceMax) {
    tr->t[level] = now_ticks();
    tr->nf[level] = nf;
    tr->m[level] = m;
    tr->mode[level] = mode;
    tr->n = level + 1;
  }
}

struct FrontierQueues {
  int* q;          // small queue
  int* mq;         // mid queue
  int* hq;         // huge queue
  unsigned* hoff;  // first chunk index of each huge entry
};

IG_DEV unsigned huge_count(unsigned long long hpack) { return static_cast<unsigned>(hpack & 0xffffffffull); }
IG_DEV unsigned huge_chunks(unsigned long long hpack) { return static_cast<unsigned>(hpack >> 32); }

IG_DEV void counters_clear(FrontierCounters* c)
{
  st_relaxed(&c->n, 0u);
  st_relaxed(&c->nm, 0u);
  st_relaxed(&c->hpack, 0ull);
  st_relaxed(&c->m, 0ull);
  st_relaxed(&c->aux, 0ull);
}

// Snapshot of a counter set (device-scope loads).
IG_DEV FrontierCounters counters_load(const FrontierCounters* c)
{
  FrontierCounters r;
  r.n = ld_relaxed(&c->n);
  r.nm = ld_relaxed(&c->nm);
  r.hpack = ld_relaxed(&c->hpack);
  r.m = ld_relaxed(&c->m);
  r.aux = ld_relaxed(&c->aux);
  return r;
}

// Wave-collective: every active lane with `pred` appends vertex u (whose list
// has deg entries) to the frontier `fq`/`c`.  All 64 lanes must be converged.
// kMid is the list length from which a vertex goes to the mid queue.
template <int kMid = kMidDeg>
IG_DEV void frontier_emit(bool pred, int u, int deg, const FrontierQueues& fq, FrontierCounters* c)
{
  const bool huge = pred && deg >= kHugeDeg;
  const bool mid = pred && deg >= kMid && !huge;
  const bool small = pred && deg < kMid;
  const unsigned long long dsum = wave_sum(static_cast<unsigned long long>(pred ? deg : 0));
  if (lane_id() == 0 && dsum != 0) atomic_add_relaxed(&c->m, dsum);
  wave_append(small, u, fq.q, &c->n);
  wave_append(mid, u, fq.mq, &c->nm);
  const unsigned long long hm = __ballot(huge);
  if (hm != 0) {
    const int chunks = huge ? (deg + kHugeChunk - 1) / kHugeChunk : 0;
    const int incl = wave_inclusive_scan(chunks);
    const int total = wave_shfl(incl, kWave - 1);
    const int leader = __ffsll(static_cast<long long>(hm)) - 1;
    unsigned long long old = 0;
    if (lane_id() == leader)
      old = atomic_add_relaxed(&c->hpack, (static_cast<unsigned long long>(total) << 32) | __popcll(hm));
    old = wave_shfl(old, leader);
    if (huge) {
      const unsigned slot = static_cast<unsigned>(old & 0xffffffffull) + lanes_below(hm);
      st_relaxed(&fq.hq[slot], u);
      st_relaxed(&fq.hoff[slot], static_cast<unsigned>(old >> 32) + static_cast<unsigned>(incl - chunks));
    }
  }
}

// Workgroup-aggregated version of frontier_emit (see emit.hpp): small and mid
// appends and the degree sum are buffered in LDS and flushed with one atomic
// per queue by fe_flush; huge entries (rare) still go straight to global.
constexpr int kEmitSmallCap = 4096;
constexpr int kEmitMidCap = 512;

template <int kSmallCap, int kMidCap>
struct FrontierEmitterT {
  BlockQueue<kSmallCap> small;
  BlockQueue<kMidCap> mid;
  BlockSum m;
};
using FrontierEmitter = FrontierEmitterT<kEmitSmallCap, kEmitMidCap>;
// For collect_heavy over a whole graph with a low threshold: no small queue,
// room for a workgroup's whole share of mid lists (an append beyond the
// buffer costs a same-address global atomic, about 10 ns each, serialised).
using HeavyCollector = FrontierEmitterT<64, 2048>;

// Workgroup-collective.
template <int S, int M>
IG_DEV void fe_reset(FrontierEmitterT<S, M>& E)
{
  if (threadIdx.x == 0) {
    E.small.count = 0;
    E.mid.count = 0;
    E.m.v = 0ull;
  }
  __syncthreads();
}

// Wave-collective.
template <int kMid = kMidDeg, int S, int M>
IG_DEV void frontier_emit_block(FrontierEmitterT<S, M>& E, bool pred, int u, int deg, const FrontierQueues& fq,
                                FrontierCounters* c)
{
  const bool huge = pred && deg >= kHugeDeg;
  const bool mid = pred && deg >= kMid && !huge;
  const bool small = pred && deg < kMid;
  bs_add(E.m, static_cast<unsigned long long>(pred ? deg : 0));
  bq_append(E.small, small, u, fq.q, &c->n);
  bq_append(E.mid, mid, u, fq.mq, &c->nm);
  const unsigned long long hm = __ballot(huge);
  if (hm != 0) {
    const int chunks = huge ? (deg + kHugeChunk - 1) / kHugeChunk : 0;
    const int incl = wave_inclusive_scan(chunks);
    const int total = wave_shfl(incl, kWave - 1);
    const int leader = __ffsll(static_cast<long long>(hm)) - 1;
    unsigned long long old = 0;
    if (lane_id() == leader)
      old = atomic_add_relaxed(&c->hpack, (static_cast<unsigned long long>(total) << 32) | __popcll(hm));
    old = wave_shfl(old, leader);
    if (huge) {
      const unsigned slot = static_cast<unsigned>(old & 0xffffffffull) + lanes_below(hm);
      st_relaxed(&fq.hq[slot], u);
      st_relaxed(&fq.hoff[slot], static_cast<unsigned>(old >> 32) + static_cast<unsigned>(incl - chunks));
    }
  }
}

// Workgroup-collective.
template <int S, int M>
IG_DEV void fe_flush(FrontierEmitterT<S, M>& E, const FrontierQueues& fq, FrontierCounters* c)
{
  bq_flush(E.small, fq.q, &c->n);
  bq_flush(E.mid, fq.mq, &c->nm);
  if (threadIdx.x == 0 && E.m.v != 0ull) atomic_add_relaxed(&c->m, E.m.v);
  __syncthreads();
}

// Index j of the huge entry owning chunk `item`: the last j with hoff[j] <= item.
IG_DEV int huge_owner(const unsigned* hoff, int nh, unsigned item)
{
  int lo = 0, hi = nh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ld_relaxed(&hoff[mid]) <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Workgroup-collective grid-stride pass over vertices 0..nodes-1 (64 per
// wave) that appends every vertex with a list of >= kMid entries to the mid /
// huge queues, aggregated per workgroup.
template <int kMid = kMidDeg, int S, int M>
IG_DEV void collect_heavy(FrontierEmitterT<S, M>& E, int nodes, const int* __restrict__ nindex,
                          const FrontierQueues& fq, FrontierCounters* c)
{
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  const int nw = (nodes + kWave - 1) / kWave;
  fe_reset(E);
  for (int w = blockIdx.x * wpb + threadIdx.x / kWave; w < nw; w += gridDim.x * wpb) {
    const int v = w * kWave + lane;
    const int deg = v < nodes ? nindex[v + 1] - nindex[v] : 0;
    frontier_emit_block<kMid>(E, deg >= kMid, v, deg, fq, c);
  }
  fe_flush(E, fq, c);
}

// ---- expansion helpers ------------------------------------------------------
// Each calls f(v, e, active) with all lanes of the wave (or all threads of the
// workgroup) converged; `active` says whether (v, e) is a real CSR entry of a
// frontier vertex v.  f may use wave-collective operations.

// Huge lists: one workgroup per kHugeChunk-entry item.
template <typename F>
IG_DEV void expand_huge(const FrontierQueues& in, unsigned nh, unsigned nch, const int* __restrict__ nindex, F&& f)
{
  __shared__ int s_owner;
  for (unsigned item = blockIdx.x; item < nch; item += gridDim.x) {
    if (threadIdx.x == 0) s_owner = huge_owner(in.hoff, static_cast<int>(nh), item);
    __syncthreads();
    const int j = s_owner;
    __syncthreads();
    const int v = ld_relaxed(&in.hq[j]);
    const int beg = nindex[v] + static_cast<int>(item - ld_relaxed(&in.hoff[j])) * kHugeChunk;
    const int end = min(nindex[v + 1], beg + kHugeChunk);
    for (int base = beg; base < end; base += blockDim.x) {
      const int e = base + static_cast<int>(threadIdx.x);
      f(v, e, e < end);
    }
  }
}

// Mid lists: one wave per vertex.
template <typename F>
IG_DEV void expand_mid(const FrontierQueues& in, unsigned nm, const int* __restrict__ nindex, F&& f)
{
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  for (unsigned i = blockIdx.x * wpb + threadIdx.x / kWave; i < nm; i += gridDim.x * wpb) {
    const int v = ld_relaxed(&in.mq[i]);
    const int beg = nindex[v], end = nindex[v + 1];
    for (int base = beg; base < end; base += kWave) {
      const int e = base + lane;
      f(v, e, e < end);
    }
  }
}

// Small lists: 64 vertices per wave, 64 CSR entries per step.
template <typename F>
IG_DEV void expand_small(const int* __restrict__ q, unsigned n, const int* __restrict__ nindex, F&& f)
{
  const int lane = lane_id();
  const int wpb = blockDim.x / kWave;
  for (unsigned c = blockIdx.x * wpb + threadIdx.x / kWave; c * kWave < n; c += gridDim.x * wpb) {
    const unsigned idx = c * kWave + lane;
    const bool valid = idx < n;
    const int v = valid ? ld_relaxed(&q[idx]) : 0;
    const int beg = valid ? nindex[v] : 0;
    const int deg = valid ? nindex[v + 1] - beg : 0;
    const int incl = wave_inclusive_scan(deg);
    const int total = wave_shfl(incl, kWave - 1);
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int ov = wave_shfl(v, o);
      const int obeg = wave_shfl(beg, o);
      const int oexcl = wave_shfl(incl - deg, o);
      f(ov, obeg + (k - oexcl), k < total);
    }
  }
}

}  // namespace indigo
