// Compiles the reference's OpenMP BFS variant named by REF_FILE with its
// harness (include/bfs_vertex_omp.h) and exposes CPUbfs_vertex with the
// thread count set through omp_set_num_threads.
#include <omp.h>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_bfs_omp(int nodes, int edges, int* nindex, int* nlist, int src, int* dist, int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  double rt = 0;
  omp_set_num_threads(threads);
  free(ref_capture_stdout([&] { rt = CPUbfs_vertex(src, g, dist); }));
  return rt;
}
