// Host-side check of the persistent-grid arithmetic (csrc/grid_plan.hpp),
// built with g++ by tests/test_grid_plan.py: per-CU residency from the
// occupancy answer, the grid under each bound, and the cap after a grid
// that did not assemble.  Prints "ok" and exits 0 when every check holds.
#include <cstdio>

#include "grid_plan.hpp"

using indigo::GridLimits;
using indigo::lowered_cap;
using indigo::plan_grid;
using indigo::plan_per_cu;

static int failures = 0;

#define CHECK_EQ(got, want)                                                                      \
  do {                                                                                           \
    const long long g_ = (got), w_ = (want);                                                     \
    if (g_ != w_) {                                                                              \
      std::fprintf(stderr, "%s:%d: %s = %lld, expected %lld\n", __FILE__, __LINE__, #got, g_, w_); \
      failures++;                                                                                \
    }                                                                                            \
  } while (0)

static GridLimits limits(int cus, int bpc, int max_blocks, int max_wg, int learned)
{
  GridLimits l;
  l.compute_units = cus;
  l.hint_blocks_per_cu = bpc;
  l.max_blocks = max_blocks;
  l.hint_max_workgroups = max_wg;
  l.learned_cap = learned;
  return l;
}

int main()
{
  // occupancy that fills every wave slot of a 2048-thread CU leaves one out
  CHECK_EQ(plan_per_cu(4, 512, 2048), 3);
  CHECK_EQ(plan_per_cu(2, 1024, 2048), 1);
  CHECK_EQ(plan_per_cu(3, 512, 2048), 3);
  CHECK_EQ(plan_per_cu(1, 1024, 1024), 1);  // a single workgroup is never dropped
  CHECK_EQ(plan_per_cu(0, 256, 2048), 0);
  // kernels held to 80 scalar registers fill the CU
  CHECK_EQ(plan_per_cu(4, 512, 2048, true), 4);
  CHECK_EQ(plan_per_cu(2, 1024, 2048, true), 2);
  CHECK_EQ(plan_per_cu(3, 512, 2048, true), 3);

  // each bound on its own, then the smallest of several
  CHECK_EQ(plan_grid(3, limits(256, 0, 0, 0, 0)), 768);
  CHECK_EQ(plan_grid(3, limits(256, 1, 0, 0, 0)), 256);   // INDIGO_HINT_BLOCKS_PER_CU
  CHECK_EQ(plan_grid(3, limits(256, 8, 0, 0, 0)), 768);   // a hint above occupancy does not raise it
  CHECK_EQ(plan_grid(3, limits(256, 0, 600, 0, 0)), 600); // the caller's bound (PageRank's max grid)
  CHECK_EQ(plan_grid(3, limits(256, 0, 0, 64, 0)), 64);   // a serving slot's share
  CHECK_EQ(plan_grid(3, limits(256, 0, 0, 0, 512)), 512); // learned cap
  CHECK_EQ(plan_grid(3, limits(256, 2, 600, 700, 400)), 400);
  CHECK_EQ(plan_grid(3, limits(256, 2, 300, 700, 400)), 300);
  CHECK_EQ(plan_grid(3, limits(256, 1, 600, 700, 400)), 256);
  CHECK_EQ(plan_grid(0, limits(256, 0, 0, 0, 0)), 1);     // never an empty grid

  // after a grid that did not assemble: one workgroup per CU fewer while the
  // grid spans several per CU, else a quarter fewer; 0 below one workgroup
  CHECK_EQ(lowered_cap(768, 256), 512);
  CHECK_EQ(lowered_cap(512, 256), 256);
  CHECK_EQ(lowered_cap(256, 256), 192);
  CHECK_EQ(lowered_cap(64, 256), 48);
  CHECK_EQ(lowered_cap(3, 256), 2);
  CHECK_EQ(lowered_cap(2, 256), 1);
  CHECK_EQ(lowered_cap(1, 256), 0);

  // a serving slot capped at 64 workgroups whose grid did not assemble
  // relaunches smaller than the share, not at the same grid, and repeated
  // failures walk down to one workgroup in a bounded number of steps
  int grid = plan_grid(3, limits(256, 0, 0, 64, 0));
  int steps = 0;
  while (grid > 1) {
    const int cap = lowered_cap(grid, 256);
    const int next = plan_grid(3, limits(256, 0, 0, 64, cap));
    CHECK_EQ(next < grid, 1);
    grid = next;
    steps++;
  }
  CHECK_EQ(steps <= 16, 1);
  // from the whole GPU, the first lowering removes one workgroup per CU
  CHECK_EQ(plan_grid(3, limits(256, 0, 0, 0, lowered_cap(768, 256))), 512);

  if (failures == 0) std::printf("ok\n");
  return failures == 0 ? 0 : 1;
}
