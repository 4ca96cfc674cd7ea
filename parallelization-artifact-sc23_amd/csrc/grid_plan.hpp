// Grid sizing of the persistent kernels, as plain arithmetic (no HIP calls),
// so that it can be checked on the host without a GPU.
//
// A persistent grid is occupancy x CUs workgroups, clamped by the caller's
// bound, by the context's hints and by what the context has learned from
// grids that did not assemble.  A grid that does not assemble within the
// residency deadline (the kernels' residency_barrier aborts it) is launched
// again at a lower cap; that cap belongs to the context and to the kernel, so
// a grid starved by work of another context (serving slots share the GPU)
// does not shrink the grids of other contexts, and it ends with the context.
#pragma once

#include <algorithm>

namespace indigo {

// Co-resident workgroups per CU from the occupancy API's answer.  A grid that
// needs every wave slot of every CU (for example 4 x 512 threads on a
// 2048-thread CU) did not assemble on MI355X for kernels of more than 80
// scalar registers, which the occupancy API does not account for, so one
// workgroup is left out; kernels held to 80 (full_cu) fill the CU.
inline int plan_per_cu(int occupancy, int block, int max_threads_per_cu, bool full_cu = false)
{
  int n = std::max(occupancy, 0);
  if (!full_cu && n > 1 && static_cast<long long>(n) * block >= max_threads_per_cu) n -= 1;
  return n;
}

struct GridLimits {
  int compute_units = 0;
  int hint_blocks_per_cu = 0;   // INDIGO_HINT_BLOCKS_PER_CU (0: none)
  int max_blocks = 0;           // the caller's bound (0: none)
  int hint_max_workgroups = 0;  // INDIGO_HINT_MAX_WORKGROUPS (0: none)
  int learned_cap = 0;          // from earlier grids that did not assemble (0: none)
};

// Workgroups of one persistent launch (at least 1).
inline int plan_grid(int per_cu, const GridLimits& l)
{
  if (l.hint_blocks_per_cu > 0) per_cu = std::min(per_cu, l.hint_blocks_per_cu);
  long long grid = static_cast<long long>(std::max(per_cu, 0)) * std::max(l.compute_units, 0);
  if (l.max_blocks > 0) grid = std::min<long long>(grid, l.max_blocks);
  if (l.hint_max_workgroups > 0) grid = std::min<long long>(grid, l.hint_max_workgroups);
  if (l.learned_cap > 0) grid = std::min<long long>(grid, l.learned_cap);
  return static_cast<int>(std::max<long long>(grid, 1));
}

// The cap after a grid of `launched` workgroups did not assemble: one
// workgroup per CU fewer while the grid spans several per CU, else a quarter
// fewer (at least one).  0 when a grid of one workgroup did not assemble:
// nothing smaller exists.
inline int lowered_cap(int launched, int compute_units)
{
  if (launched <= 1) return 0;
  if (compute_units > 0 && launched > compute_units) return launched - compute_units;
  return std::max(1, launched - std::max(1, launched / 4));
}

}  // namespace indigo
