// Compiles the reference's C++ CC variant named by REF_FILE with its harness
// (include/cc_vertex_cpp.h) and exposes CPUcc_vertex.
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_cc_threads(int nodes, int edges, int* nindex, int* nlist, int* label, int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  double rt = 0;
  free(ref_capture_stdout([&] { rt = CPUcc_vertex(g, label, threads); }));
  return rt;
}
