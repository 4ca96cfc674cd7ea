// Compiles the reference's OpenMP edge-centric CC variant named by REF_FILE
// with its harness (include/cc_edge_omp.h) and exposes CPUcc_edge with the
// thread count set through omp_set_num_threads.  The edge-source array is
// built as the harness builds it (cc_edge_omp.h), outside the timed call.
#include <omp.h>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_cc_omp(int nodes, int edges, int* nindex, int* nlist, int* label, int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  int* sp = new int[edges > 0 ? edges : 1];
  for (int v = 0; v < nodes; v++)
    for (int j = nindex[v]; j < nindex[v + 1]; j++) sp[j] = v;
  double rt = 0;
  omp_set_num_threads(threads);
  free(ref_capture_stdout([&] { rt = CPUcc_edge(g, label, sp); }));
  delete[] sp;
  return rt;
}
