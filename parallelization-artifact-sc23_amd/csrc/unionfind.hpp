// Concurrent union-find on a parent array, shared by the CC kernels.
//
// Every link hooks the LARGER root under the smaller one with a CAS, so
// parent[x] <= x always holds and each final root is the minimum id of its
// tree - once all edges are linked, the minimum id of its component, which
// is the reference's CC label (min-label propagation to a fixed point,
// e.g. codes/cuda/cc-cuda/CC_CUDA_V_Topo_Push_NonDeterm_IntType_
// ReadModifyWrite_NonPersist_Atomic_Thread_NoBoundsBug2_NoBoundsBug1.cu:47-76).
// Finds use path halving with relaxed stores: every store replaces a parent
// by one of its ancestors, so concurrent finds and links stay valid.
#pragma once

#include "device_common.hpp"

namespace indigo {

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
    // hi stopped being a root meanwhile: climb from where it now points
    ra = find_root(parent, ra == hi ? expected : ra);
    rb = find_root(parent, rb == hi ? expected : rb);
  }
}

}  // namespace indigo
