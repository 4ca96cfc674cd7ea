// Compiles the reference's C++ SSSP variant named by REF_FILE with its
// harness (include/sssp_vertex_cpp.h) and exposes CPUsssp_vertex and the
// serial oracle CPUserialDijkstra (sssp_vertex_cpp.h:88).
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_sssp_threads(int nodes, int edges, int* nindex, int* nlist, int* eweight, int src, int* dist,
                                   int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, eweight};
  double rt = 0;
  free(ref_capture_stdout([&] { rt = CPUsssp_vertex(src, g, dist, threads); }));
  return rt;
}

extern "C" void ref_sssp_serial(int nodes, int edges, int* nindex, int* nlist, int* eweight, int src, int* dist)
{
  ECLgraph g{nodes, edges, nindex, nlist, eweight};
  CPUserialDijkstra(src, g, dist);
}
