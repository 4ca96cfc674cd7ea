// Compiles the reference's C++ PageRank variant named by REF_FILE with its
// harness (include/pr_cpp.h) and exposes PR_CPU; the iteration count, which
// the reference only prints ("CPU iterations = %d."), is parsed from stdout.
#include <cstring>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_pr_threads(int nodes, int edges, int* nindex, int* nlist, float* scores, int* iterations,
                                 int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  int* degree = new int[nodes];
  for (int v = 0; v < nodes; v++) degree[v] = nindex[v + 1] - nindex[v];
  const float init = 1.0f / static_cast<float>(nodes);
  for (int v = 0; v < nodes; v++) scores[v] = init;
  double rt = 0;
  char* out = ref_capture_stdout([&] { rt = PR_CPU(g, scores, degree, threads); });
  const char* p = strstr(out, "iterations = ");
  *iterations = p ? atoi(p + strlen("iterations = ")) : -1;
  free(out);
  delete[] degree;
  return rt;
}
