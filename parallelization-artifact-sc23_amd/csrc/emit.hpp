// Workgroup-aggregated queue appends.
//
// Appending with one global atomic per wave serialises on the counter's
// address once thousands of waves emit in the same level (measured: several
// ms per level on a 4M-vertex frontier).  Here waves append into an LDS
// buffer with LDS atomics; at the end of the phase the workgroup reserves its
// global range with ONE atomic per queue and copies the buffer out
// coalesced.  A buffer that fills up spills the overflowing part of a wave's
// batch straight to the global queue, so any number of appends is correct.
//
// Usage (all calls workgroup-uniform except append, which is wave-collective):
//   bq_reset(B);  ... bq_append(B, pred, item, gq, gcount) ...;  bq_flush(B, gq, gcount);
// Global queue contents are written with device-scope stores.
#pragma once

#include "device_common.hpp"

namespace indigo {

template <int kCap>
struct BlockQueue {
  int count;  // appends so far this phase (may exceed kCap: the excess went global)
  int base;   // global slot of buf[0] after the flush reservation
  int buf[kCap];
};

template <int kCap>
IG_DEV void bq_reset(BlockQueue<kCap>& B)
{
  if (threadIdx.x == 0) B.count = 0;
  __syncthreads();
}

// Wave-collective: lanes with `pred` append `item`.
template <int kCap, typename C>
IG_DEV void bq_append(BlockQueue<kCap>& B, bool pred, int item, int* gq, C* gcount)
{
  const unsigned long long m = __ballot(pred);
  if (m == 0) return;
  const int lane = lane_id();
  const int leader = __ffsll(static_cast<long long>(m)) - 1;
  const int pop = __popcll(m);
  int slot = 0;
  if (lane == leader) slot = atomicAdd(&B.count, pop);
  slot = wave_shfl(slot, leader);
  const int idx = slot + lanes_below(m);
  if (slot + pop <= kCap) {
    if (pred) B.buf[idx] = item;
    return;
  }
  // spill: the part of this batch at or beyond kCap goes to the global queue
  const int first_spill = slot > kCap ? slot : kCap;
  const int spill = slot + pop - first_spill;
  C gbase = 0;
  if (lane == leader) gbase = atomic_add_relaxed(gcount, static_cast<C>(spill));
  gbase = wave_shfl(gbase, leader);
  if (pred) {
    if (idx < kCap) B.buf[idx] = item;
    else st_relaxed(&gq[static_cast<int>(gbase) + (idx - first_spill)], item);
  }
}

// Workgroup-collective: reserves this workgroup's range with one atomic and
// copies the buffered items out.
template <int kCap, typename C>
IG_DEV void bq_flush(BlockQueue<kCap>& B, int* gq, C* gcount)
{
  __syncthreads();
  const int n = B.count < kCap ? B.count : kCap;
  if (threadIdx.x == 0 && n > 0) B.base = static_cast<int>(atomic_add_relaxed(gcount, static_cast<C>(n)));
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) st_relaxed(&gq[B.base + i], B.buf[i]);
  __syncthreads();
}

// Workgroup-level 64-bit sum flushed with one atomic.
struct BlockSum {
  unsigned long long v;
};

IG_DEV void bs_reset(BlockSum& S)
{
  if (threadIdx.x == 0) S.v = 0ull;
  __syncthreads();
}

// Wave-collective: adds the wave's total of x.
IG_DEV void bs_add(BlockSum& S, unsigned long long x)
{
  x = wave_sum(x);
  if (lane_id() == 0 && x != 0ull) atomicAdd(&S.v, x);
}

IG_DEV void bs_flush(BlockSum& S, unsigned long long* g)
{
  __syncthreads();
  if (threadIdx.x == 0 && S.v != 0ull) atomic_add_relaxed(g, S.v);
  __syncthreads();
}

}  // namespace indigo
