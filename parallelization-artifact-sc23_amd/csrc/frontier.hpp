// Degree-binned frontier queues for the persistent traversal kernels.
//
// A frontier is two queues:
//   * normal: vertex ids with degree < kHugeDeg, expanded by 64-lane waves with
//     an in-wave load balancer (prefix sum of degrees + owner search), so a
//     wave always works on 64 consecutive CSR entries;
//   * huge:   vertex ids with degree >= kHugeDeg, split into kHugeChunk-edge
//     work items spread over all workgroups.  Its counter packs
//     (chunk count << 32 | vertex count) so one 64-bit atomic allocates both
//     the slot and the vertex's first chunk index; hoff[] is therefore sorted
//     by slot and a chunk's owner is found by binary search.
// Appends are wave-aggregated (ballot + mbcnt prefix, one atomic per wave).
// Counters are triple-buffered across levels so that the set written in level
// L, read in L+1 and cleared in L+2 never has two roles at once.
#pragma once

#include "device_common.hpp"

namespace indigo {

constexpr int kHugeDeg = 4096;
constexpr int kHugeChunk = 2048;

struct alignas(128) FrontierCounters {
  unsigned int n;            // normal queue length
  unsigned int pad0;
  unsigned long long hpack;  // (huge chunks << 32) | huge queue length
  unsigned long long m;      // sum of degrees over the frontier
  unsigned long long aux;    // algorithm-specific
  unsigned long long pad[12];
};

struct FrontierQueues {
  int* q;          // normal queue
  int* hq;         // huge queue
  unsigned* hoff;  // first chunk index of each huge entry
};

IG_DEV unsigned huge_count(unsigned long long hpack) { return static_cast<unsigned>(hpack & 0xffffffffull); }
IG_DEV unsigned huge_chunks(unsigned long long hpack) { return static_cast<unsigned>(hpack >> 32); }

// Wave-collective: every active lane with `pred` appends vertex u (whose CSR
// range starts at beg and has deg entries) to the frontier `fq`/`c`.
// All 64 lanes must be converged at the call.
IG_DEV void frontier_emit(bool pred, int u, int deg, const FrontierQueues& fq, FrontierCounters* c)
{
  const bool huge = pred && deg >= kHugeDeg;
  const bool norm = pred && !huge;
  const unsigned long long dsum = wave_sum(static_cast<unsigned long long>(pred ? deg : 0));
  if (lane_id() == 0 && dsum != 0) atomic_add_relaxed(&c->m, dsum);
  wave_append(norm, u, fq.q, &c->n);
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
      fq.hq[slot] = u;
      fq.hoff[slot] = static_cast<unsigned>(old >> 32) + static_cast<unsigned>(incl - chunks);
    }
  }
}

// Index j of the huge entry owning chunk `item`: the last j with hoff[j] <= item.
IG_DEV int huge_owner(const unsigned* hoff, int nh, unsigned item)
{
  int lo = 0, hi = nh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (hoff[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace indigo
