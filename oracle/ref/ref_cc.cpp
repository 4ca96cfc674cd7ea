// Compiles the reference's C++ CC variant named by REF_FILE with its harness
// and exposes its threaded entry: CPUcc_vertex (include/cc_vertex_cpp.h), or
// with REF_EDGE CPUcc_edge (include/cc_edge_cpp.h), whose edge-source array
// is built as that harness builds it, outside the timed call.
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_cc_threads(int nodes, int edges, int* nindex, int* nlist, int* label, int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  double rt = 0;
#ifdef REF_EDGE
  int* sp = new int[edges > 0 ? edges : 1];
  for (int v = 0; v < nodes; v++)
    for (int j = nindex[v]; j < nindex[v + 1]; j++) sp[j] = v;
  free(ref_capture_stdout([&] { rt = CPUcc_edge(g, label, sp, threads); }));
  delete[] sp;
#else
  free(ref_capture_stdout([&] { rt = CPUcc_vertex(g, label, threads); }));
#endif
  return rt;
}
