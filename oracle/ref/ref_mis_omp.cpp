// Compiles the reference's OpenMP MIS variant named by REF_FILE with its
// harness (include/mis_vertex_omp.h) and exposes OMPmis_vertex with the
// thread count set through omp_set_num_threads (the harness itself reads
// OMP_NUM_THREADS); status codes are widened from unsigned char to int
// (1 included, 2 excluded).
#include <omp.h>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_mis_omp(int nodes, int edges, int* nindex, int* nlist, unsigned* priority, int* status,
                              int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  unsigned char* st = new unsigned char[nodes > 0 ? nodes : 1];
  data_type* pr = new data_type[nodes > 0 ? nodes : 1];
  double rt = 0;
  omp_set_num_threads(threads);
  free(ref_capture_stdout([&] { rt = OMPmis_vertex(g, pr, st); }));
  for (int v = 0; v < nodes; v++) {
    status[v] = st[v];
    priority[v] = static_cast<unsigned>(pr[v]);
  }
  delete[] st;
  delete[] pr;
  return rt;
}
