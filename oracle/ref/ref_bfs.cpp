// Compiles the reference's C++ BFS variant named by REF_FILE together with
// its harness (include/bfs_vertex_cpp.h) and exposes the threaded entry
// CPPbfs_vertex and the serial oracle CPUserialDijkstra (bfs_vertex_cpp.h:52).
#include "ref_common.hpp"
#define main indigo_ref_unused_main
#include REF_FILE
#undef main

static ECLgraph ref_graph(int nodes, int edges, int* nindex, int* nlist, int* eweight)
{
  ECLgraph g;
  g.nodes = nodes;
  g.edges = edges;
  g.nindex = nindex;
  g.nlist = nlist;
  g.eweight = eweight;
  return g;
}

extern "C" double ref_bfs_threads(int nodes, int edges, int* nindex, int* nlist, int src, int* dist, int threads)
{
  const ECLgraph g = ref_graph(nodes, edges, nindex, nlist, nullptr);
  double rt = 0;
  free(ref_capture_stdout([&] { rt = CPPbfs_vertex(src, g, dist, threads); }));
  return rt;
}

extern "C" void ref_bfs_serial(int nodes, int edges, int* nindex, int* nlist, int src, int* dist)
{
  const ECLgraph g = ref_graph(nodes, edges, nindex, nlist, nullptr);
  CPUserialDijkstra(src, g, dist);
}
