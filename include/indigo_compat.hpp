// indigo_compat.hpp - the reference harness's entry functions, with their
// signatures and behaviour, implemented over the C ABI in indigo_hip.h.
//
// The reference declares one static entry function per harness header
// (include/{alg}_cuda.h); the including translation unit typedefs basic_t,
// data_type and score_type first.  Here each entry is a function template on
// those types, so a harness that calls GPUbfs_vertex(src, d_g, dist) compiles
// unchanged.  Failures print the library's message and exit(-1), as the
// reference's CheckCuda does.  The graph struct handed in has device pointers
// (allocated with indigo_malloc / indigo_memcpy_h2d) and host nodes/edges,
// exactly as the reference harness prepares d_g.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "ECLgraph.h"
#include "indigo_hip.h"

namespace indigo_compat {

inline void check(int rc, const char* what)
{
  if (rc != INDIGO_OK) {
    fprintf(stderr, "ERROR: %s failed: %s\n", what, indigo_last_error());
    exit(-1);
  }
}

template <typename T>
inline int* as_int(T* p)
{
  static_assert(sizeof(T) == sizeof(int) && std::is_integral<T>::value, "outputs are 32-bit integers");
  return reinterpret_cast<int*>(p);
}

}  // namespace indigo_compat

// GPUinfo, include/bfs_vertex_cuda.h:63-73: prints the device and returns the
// persistent-grid thread bound (compute units x max threads per CU).
static inline int GPUinfo(const int d)
{
  int cus = 0, tpcu = 0;
  char name[256] = {0};
  indigo_compat::check(indigo_device_info(d, &cus, &tpcu, name, sizeof(name)), "device query");
  printf("GPU: %s with %d CUs and %d mTpCU\n", name, cus, tpcu);
  return cus * tpcu;
}

// GPUbfs_vertex, include/bfs_vertex_cuda.h:61.
template <typename basic_t>
static double GPUbfs_vertex(const int src, const ECLgraph& g, basic_t* const dist)
{
  double s = 0.0;
  indigo_compat::check(indigo_bfs_vertex(src, g.nodes, g.edges, g.nindex, g.nlist, indigo_compat::as_int(dist), &s),
                       "GPUbfs_vertex");
  return s;
}

// GPUbfs_edge, include/bfs_edge_cuda.h:61 (sp: host edge sources).
template <typename basic_t>
static double GPUbfs_edge(const int src, const ECLgraph& g, basic_t* const dist, const int* const sp)
{
  double s = 0.0;
  indigo_compat::check(
      indigo_bfs_edge(src, g.nodes, g.edges, g.nindex, g.nlist, sp, indigo_compat::as_int(dist), &s), "GPUbfs_edge");
  return s;
}

// GPUsssp_vertex, include/sssp_vertex_cuda.h:60.
template <typename basic_t>
static double GPUsssp_vertex(const int src, const ECLgraph& g, basic_t* const dist)
{
  double s = 0.0;
  indigo_compat::check(indigo_sssp_vertex(src, g.nodes, g.edges, g.nindex, g.nlist, g.eweight,
                                          indigo_compat::as_int(dist), &s),
                       "GPUsssp_vertex");
  return s;
}

// GPUsssp_edge, include/sssp_edge_cuda.h:60 (sp: host edge sources).
template <typename basic_t>
static double GPUsssp_edge(const int src, const ECLgraph& g, basic_t* const dist, const int* const sp)
{
  double s = 0.0;
  indigo_compat::check(indigo_sssp_edge(src, g.nodes, g.edges, g.nindex, g.nlist, g.eweight, sp,
                                        indigo_compat::as_int(dist), &s),
                       "GPUsssp_edge");
  return s;
}

// GPUcc_vertex, include/cc_vertex_cuda.h:57: label = smallest vertex id of
// the component.
template <typename basic_t>
static double GPUcc_vertex(const ECLgraph& g, basic_t* const label)
{
  double s = 0.0;
  indigo_compat::check(indigo_cc_vertex(g.nodes, g.edges, g.nindex, g.nlist, indigo_compat::as_int(label), &s),
                       "GPUcc_vertex");
  return s;
}

// GPUcc_edge, include/cc_edge_cuda.h:57 (sp: host edge sources).
template <typename basic_t>
static double GPUcc_edge(const ECLgraph& g, basic_t* const label, const int* const sp)
{
  double s = 0.0;
  indigo_compat::check(indigo_cc_edge(g.nodes, g.edges, g.nindex, g.nlist, sp, indigo_compat::as_int(label), &s),
                       "GPUcc_edge");
  return s;
}

// GPUmis_vertex, include/mis_vertex_cuda.h:54.  A signed data_type selects the
// CUDA variants' signed priority order, unsigned the C++ order.
template <typename data_type>
static double GPUmis_vertex(const ECLgraph& g, data_type* const priority, int* const status)
{
  static_assert(sizeof(data_type) == sizeof(unsigned), "32-bit priorities");
  double s = 0.0;
  indigo_compat::check(indigo_mis_vertex(g.nodes, g.edges, g.nindex, g.nlist, std::is_signed<data_type>::value ? 1 : 0,
                                         reinterpret_cast<unsigned*>(priority), status, &s),
                       "GPUmis_vertex");
  return s;
}

// GPUmis_edge, include/mis_edge_cuda.h:54 (d_sp: DEVICE edge sources).
template <typename data_type>
static double GPUmis_edge(const ECLgraph& g, const int* const d_sp, data_type* const priority, int* const status)
{
  static_assert(sizeof(data_type) == sizeof(unsigned), "32-bit priorities");
  double s = 0.0;
  indigo_compat::check(indigo_mis_edge(g.nodes, g.edges, g.nindex, g.nlist, d_sp, std::is_signed<data_type>::value ? 1 : 0,
                                       reinterpret_cast<unsigned*>(priority), status, &s),
                       "GPUmis_edge");
  return s;
}

// PR_GPU, include/pr_cuda.h:48: HOST graph, scores in (1/N) and out; prints
// the iteration count, the runtime and the Throughput line itself.
template <typename score_type>
static void PR_GPU(const ECLgraph g, score_type* scores, int* degree)
{
  static_assert(std::is_same<score_type, float>::value, "PageRank scores are FP32");
  double s = 0.0;
  int iterations = 0;
  indigo_compat::check(indigo_pr(g.nodes, g.edges, g.nindex, g.nlist, degree, scores, &iterations, &s), "PR_GPU");
  printf("GPU iterations = %d.\n", iterations);
  printf("GPU runtime: %.6fs\n\n", s);
  printf("Throughput: %.6f gigaedges/s\n", 0.000000001 * g.edges / s);
}

// GPUtc_vertex, include/tc_vertex_cuda.h:42: count of triangles (64-bit
// inside; narrowed to basic_t as the caller declared it).
template <typename basic_t>
static double GPUtc_vertex(basic_t& count, const int nodes, const int* const nindex, const int* const nlist)
{
  double s = 0.0;
  long long c = 0;
  indigo_compat::check(indigo_tc_vertex(&c, nodes, nindex, nlist, &s), "GPUtc_vertex");
  count = static_cast<basic_t>(c);
  return s;
}

// GPUtc_edge, include/tc_edge_cuda.h:42 (sp: DEVICE edge sources).
template <typename basic_t>
static double GPUtc_edge(basic_t& count, const int edges, const int* const nindex, const int* const nlist,
                         const int* const sp)
{
  double s = 0.0;
  long long c = 0;
  indigo_compat::check(indigo_tc_edge(&c, edges, nindex, nlist, sp, &s), "GPUtc_edge");
  count = static_cast<basic_t>(c);
  return s;
}
