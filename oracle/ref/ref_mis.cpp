// Compiles the reference's C++ MIS variant named by REF_FILE with its harness
// (include/mis_vertex_cpp.h) and exposes CPPmis_vertex; status codes are
// widened from unsigned char to int (1 included, 2 excluded).
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_mis_threads(int nodes, int edges, int* nindex, int* nlist, unsigned* priority, int* status,
                                  int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  unsigned char* st = new unsigned char[nodes > 0 ? nodes : 1];
  data_type* pr = new data_type[nodes > 0 ? nodes : 1];
  double rt = 0;
  free(ref_capture_stdout([&] { rt = CPPmis_vertex(g, pr, st, threads); }));
  for (int v = 0; v < nodes; v++) {
    status[v] = st[v];
    priority[v] = static_cast<unsigned>(pr[v]);
  }
  delete[] st;
  delete[] pr;
  return rt;
}
