// Compiles the reference's OpenMP TC variant named by REF_FILE with its
// harness (include/tc_vertex_omp.h) and exposes CPUtc_vertex with the thread
// count set through omp_set_num_threads (the harness itself reads
// OMP_NUM_THREADS).  The count is the reference's int.
#include <omp.h>

#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

extern "C" double ref_tc_omp(int nodes, int* nindex, int* nlist, long long* count, int threads)
{
  basic_t c = 0;
  double rt = 0;
  omp_set_num_threads(threads);
  free(ref_capture_stdout([&] { rt = CPUtc_vertex(c, nodes, nindex, nlist); }));
  *count = static_cast<long long>(c);
  return rt;
}
