// Drop-in harness: BFS, edge form.  Same argv
// (input_file_name source_node_number verify) and stdout contract as the
// reference's include/bfs_edge_cuda.h main(): one "Throughput: <GTEPS> gigaedges/s" line, the
// median of 3 timed calls of GPUbfs_edge.
#include "common.hpp"

typedef int basic_t;

int main(int argc, char* argv[])
{
  printf("bfs edge-centric HIP (%s)\n", __FILE__);
  if (argc != 4) {
    fprintf(stderr, "USAGE: %s input_file_name source_node_number verify\n", argv[0]);
    exit(-1);
  }
  ECLgraph g = readECLgraph(argv[1]);
  printf("input: %s\n", argv[1]);
  printf("nodes: %d\n", g.nodes);
  printf("edges: %d\n", g.edges);
  const int source = atoi(argv[2]);
  if (source < 0 || source >= g.nodes) {
    fprintf(stderr, "ERROR: source_node_number must be between 0 and %d\n", g.nodes);
    exit(-1);
  }
  printf("source: %d\n", source);
  const int runveri = harness::verify_flag(argv[3]);
  GPUinfo(0);
  std::vector<basic_t> distance(static_cast<size_t>(g.nodes));
  ECLgraph d_g = harness::upload(g, false);
  const std::vector<int> sp = harness::edge_sources(g);
  const double med = harness::timed_runs(3, [&] { return GPUbfs_edge(source, d_g, distance.data(), sp.data()); });
  harness::report(med, g.edges);
  int maxnode = 0;
  for (int v = 1; v < g.nodes; v++)
    if (distance[maxnode] < distance[v]) maxnode = v;
  printf("vertex %d has maximum distance %d\n", maxnode, distance[maxnode]);
  if (runveri) harness::verify_paths(g, source, distance.data(), false);
  else printf("turn off verification\n\n");
  harness::release(d_g);
  freeECLgraph(g);
  return 0;
}
