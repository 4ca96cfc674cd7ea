// Compiles the reference's C++ TC variant named by REF_FILE with its harness
// (include/tc_vertex_cpp.h) and exposes CPUtc_vertex and the serial oracle
// h_triCounting (tc_vertex_cpp.h:88).  Both count in the reference's int.
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_tc_threads(int nodes, int* nindex, int* nlist, long long* count, int threads)
{
  basic_t c = 0;
  double rt = 0;
  free(ref_capture_stdout([&] { rt = CPUtc_vertex(c, nodes, nindex, nlist, threads); }));
  *count = static_cast<long long>(c);
  return rt;
}

extern "C" long long ref_tc_serial(int nodes, int* nindex, int* nlist)
{
  return static_cast<long long>(h_triCounting(nodes, nindex, nlist));
}
