// Compiles the reference's OpenMP edge-centric SSSP variant named by REF_FILE
// with its harness (include/sssp_edge_omp.h) and exposes CPUsssp_edge with the
// thread count set through omp_set_num_threads.  The edge-source array is
// built as the harness builds it (sssp_edge_omp.h), outside the timed call.
#include <omp.h>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_sssp_omp(int nodes, int edges, int* nindex, int* nlist, int* eweight, int src, int* dist,
                               int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, eweight};
  int* sp = new int[edges > 0 ? edges : 1];
  for (int v = 0; v < nodes; v++)
    for (int j = nindex[v]; j < nindex[v + 1]; j++) sp[j] = v;
  double rt = 0;
  omp_set_num_threads(threads);
  free(ref_capture_stdout([&] { rt = CPUsssp_edge(src, g, dist, sp); }));
  delete[] sp;
  return rt;
}
