// Device-side building blocks shared by the gfx950 graph kernels:
// relaxed agent-scope atomics, 64-lane wave scans and aggregated appends,
// and a bounded grid barrier for persistent (co-resident) launches.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

namespace indigo {

constexpr int kWave = 64;
constexpr int kInf = INT_MAX;  // unreachable marker (reference bfs_vertex_cuda.h:46)

#define IG_DEV __device__ __forceinline__

// Caps a kernel at 80 scalar registers.  On MI355X a wave that holds more
// (96 or 112 after the allocation granule of 16) keeps its SIMD below 8
// resident waves, which neither the compiler's occupancy figure nor the
// occupancy API shows: a TC counting kernel ran 13.3 ms at 96 SGPRs and
// 10.7 ms at 80 with otherwise the same code, and grids of 4 x 512 threads
// per CU assemble only for kernels within 80.  The spilled values live in
// VGPR lanes.
#define IG_SGPR_CAP __attribute__((amdgpu_num_sgpr(80)))

// ---- relaxed atomics on plain int arrays (the suite's atomicRead/atomicWrite,
// include/bfs_vertex_cuda.h:49-59, as agent-scope relaxed HIP atomics) ------
template <typename T>
IG_DEV T ld_relaxed(const T* p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
IG_DEV void st_relaxed(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
IG_DEV T atomic_min_relaxed(T* p, T v)
{
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
IG_DEV T atomic_max_relaxed(T* p, T v)
{
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
IG_DEV T atomic_add_relaxed(T* p, T v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
IG_DEV bool cas_relaxed(T* p, T expected, T desired)
{
  return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

// ---- wave helpers ---------------------------------------------------------
IG_DEV int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` in lanes below the calling lane.
IG_DEV int lanes_below(unsigned long long mask)
{
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

template <typename T>
IG_DEV T wave_shfl(T v, int src_lane)
{
  return __shfl(v, src_lane, kWave);
}

// Value of v in lane `lane`, which must be the same in all lanes (readlane:
// no LDS traffic).
IG_DEV int wave_bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Inclusive prefix sum over the 64 lanes (all lanes must be active), with DPP
// row operations instead of LDS permutes: Hillis-Steele within each row of 16
// lanes (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into
// row 3 (row_bcast:15, rows 1 and 3), then lane 31's total into rows 2 and 3
// (row_bcast:31).  A lane whose DPP source is out of its row, or whose row is
// masked off, adds the `old` operand, 0.
IG_DEV int wave_inclusive_scan(int x)
{
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

template <typename T>
IG_DEV T wave_sum(T x)
{
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

template <typename T>
IG_DEV T wave_max(T x)
{
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const T y = __shfl_xor(x, d, kWave);
    x = y > x ? y : x;
  }
  return x;
}

template <typename T>
IG_DEV T wave_min(T x)
{
#pragma unroll
  for (int d = kWave / 2; d > 0; d >>= 1) {
    const T y = __shfl_xor(x, d, kWave);
    x = y < x ? y : x;
  }
  return x;
}

// Appends `item` for every lane with `pred` to queue[] using one atomic per
// wave on *counter; returns the slot (or -1).  Lanes that are inactive at the
// call simply do not take part.
template <typename C>
IG_DEV int wave_append(bool pred, int item, int* queue, C* counter)
{
  const unsigned long long m = __ballot(pred);
  if (m == 0) return -1;
  const int leader = __ffsll(static_cast<long long>(m)) - 1;
  C base = 0;
  if (lane_id() == leader) base = atomic_add_relaxed(counter, static_cast<C>(__popcll(m)));
  base = wave_shfl(base, leader);
  if (!pred) return -1;
  const int slot = static_cast<int>(base) + lanes_below(m);
  st_relaxed(&queue[slot], item);  // device-scope: readable after a coherent barrier
  return slot;
}

// Owner search for the wave load balancer: given the non-decreasing inclusive
// prefix `incl` held by lanes 0..63, returns the number of lanes whose incl <= k,
// i.e. the lane whose [incl - deg, incl) range contains k.
IG_DEV int wave_owner(int incl, int k)
{
  int lo = 0;
#pragma unroll
  for (int step = 32; step > 0; step >>= 1) {
    const int probe = lo + step - 1;
    const int v = __shfl(incl, probe < kWave ? probe : kWave - 1, kWave);
    if (probe < kWave && v <= k) lo += step;
  }
  return lo;
}

// ---- bounded grid barrier -------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs, so workgroup b
// arrives on counter b % 8 first; the last arrival of each group bumps the
// top counter and the very last one advances the generation.  Every wait is
// bounded by a wall-clock deadline: on expiry the waiter raises `abort` and
// all barriers from then on return false, so every wave leaves the kernel.
//
// Two memory-ordering modes:
//   * fenced (default): thread 0 brackets the arrival with agent-scope
//     release / acquire fences.  On gfx950 these write back / invalidate the
//     XCD's L2 (buffer_wbl2 / buffer_inv), publishing plain stores; measured
//     at 5-27 us per barrier depending on grid size and dirty data.
//   * coherent: for kernels whose data shared between workgroups is only
//     accessed with device-scope atomic operations (ld_relaxed / st_relaxed
//     / RMWs, which bypass the non-coherent caches).  Completion of those
//     accesses (s_waitcnt 0 in every wave before the workgroup barrier) is
//     then all the ordering a barrier needs: about 2.5 us.
// Waiters back off exponentially so that long waits (one-workgroup phases)
// do not flood the memory channel of the generation counter.
struct alignas(128) PaddedU32 {
  unsigned int v;
  unsigned int pad[31];
};

struct GridBarrier {
  PaddedU32 group[8];
  PaddedU32 top;
  PaddedU32 gen;
  PaddedU32 abort;
};

IG_DEV long long now_ticks() { return wall_clock64(); }

// abort codes stored in GridBarrier::abort
constexpr unsigned kAbortTimeout = 1u;      // a bounded wait expired mid-run
constexpr unsigned kAbortNotResident = 2u;  // the grid never assembled: not co-resident

// Waits until every memory operation issued by this wave has completed.
IG_DEV void wave_drain() { __builtin_amdgcn_s_waitcnt(0); }

template <bool kCoherent = false>
IG_DEV bool grid_barrier(GridBarrier* gb, long long deadline, unsigned abort_code = kAbortTimeout)
{
  // fenced: acq_rel arrivals and a release generation store chain every
  // workgroup's release to every waiter's acquire fence; coherent: relaxed
  // device-scope atomics, each completed (s_waitcnt) before the next step.
  constexpr int kRmw = kCoherent ? __ATOMIC_RELAXED : __ATOMIC_ACQ_REL;
  constexpr int kPub = kCoherent ? __ATOMIC_RELAXED : __ATOMIC_RELEASE;
  __shared__ int s_ok;
  if (kCoherent) wave_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned nblocks = gridDim.x;
    const unsigned g = blockIdx.x & 7u;
    const unsigned groups = nblocks < 8u ? nblocks : 8u;
    const unsigned gsize = (nblocks - g + 7u) >> 3;
    const unsigned gen0 = __hip_atomic_load(&gb->gen.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool last = false;
    const unsigned a = __hip_atomic_fetch_add(&gb->group[g].v, 1u, kRmw, __HIP_MEMORY_SCOPE_AGENT);
    if (a == gsize - 1u) {
      __hip_atomic_store(&gb->group[g].v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (kCoherent) wave_drain();
      const unsigned t = __hip_atomic_fetch_add(&gb->top.v, 1u, kRmw, __HIP_MEMORY_SCOPE_AGENT);
      if (t == groups - 1u) {
        __hip_atomic_store(&gb->top.v, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (kCoherent) wave_drain();
        __hip_atomic_store(&gb->gen.v, gen0 + 1u, kPub, __HIP_MEMORY_SCOPE_AGENT);
        last = true;
      }
    }
    if (!last) {
      int backoff = 1;
      while (__hip_atomic_load(&gb->gen.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen0) {
        if (__hip_atomic_load(&gb->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
          ok = 0;
          break;
        }
        if (now_ticks() > deadline) {
          unsigned expected = 0u;
          __hip_atomic_compare_exchange_strong(&gb->abort.v, &expected, abort_code, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        for (int k = 0; k < backoff; k++) __builtin_amdgcn_s_sleep(2);
        backoff = backoff < 64 ? backoff * 2 : 64;
      }
    }
    if (__hip_atomic_load(&gb->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) ok = 0;
    if (!kCoherent) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// First action of every persistent kernel: all workgroups must arrive
// within `ticks` (they are launched together; the only way to miss is a
// grid that cannot be co-resident).  On failure the host relaunches with
// fewer workgroups per CU; nothing has been written yet.
IG_DEV bool residency_barrier(GridBarrier* gb, long long ticks)
{
  return grid_barrier<true>(gb, now_ticks() + ticks, kAbortNotResident);
}

// Same as grid_barrier's abort test, for loops that run between barriers
// (one-workgroup phases): true when the kernel must stop.
IG_DEV bool kernel_aborted(GridBarrier* gb, long long deadline)
{
  if (__hip_atomic_load(&gb->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return true;
  if (now_ticks() > deadline) {
    __hip_atomic_store(&gb->abort.v, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  return false;
}

// MIS priority: the suite's integer hash (include/mis_vertex_cuda.h:174-179,
// xor-shift-multiply by 0x45d9f3b twice) of v + 712313887 (init kernels, e.g.
// MIS_CUDA_V_Topo_Push_Determ_IntType_NonPersist_Atomic_Thread.cu:51).
__host__ __device__ inline unsigned int mis_hash(unsigned int val)
{
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  return (val >> 16) ^ val;
}

}  // namespace indigo
