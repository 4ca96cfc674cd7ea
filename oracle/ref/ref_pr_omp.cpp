// Compiles the reference's OpenMP PageRank variant named by REF_FILE with its
// harness (include/pr_omp.h) and exposes PR_CPU with the thread count set
// through omp_set_num_threads.  PR_CPU returns nothing: its runtime and
// iteration count, which it prints ("CPU runtime: %.6fs", "CPU iterations =
// %d."), are parsed from stdout.
#include <omp.h>

#include <cstring>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_pr_omp(int nodes, int edges, int* nindex, int* nlist, float* scores, int* iterations,
                             int threads)
{
  ECLgraph g{nodes, edges, nindex, nlist, nullptr};
  int* degree = new int[nodes > 0 ? nodes : 1];
  for (int v = 0; v < nodes; v++) degree[v] = nindex[v + 1] - nindex[v];
  const float init = 1.0f / static_cast<float>(nodes);
  for (int v = 0; v < nodes; v++) scores[v] = init;
  omp_set_num_threads(threads);
  char* out = ref_capture_stdout([&] { PR_CPU(g, scores, degree); });
  const char* it = strstr(out, "iterations = ");
  const char* rt = strstr(out, "runtime: ");
  *iterations = it ? atoi(it + strlen("iterations = ")) : -1;
  const double seconds = rt ? atof(rt + strlen("runtime: ")) : -1.0;
  free(out);
  delete[] degree;
  return seconds;
}
