// Drop-in harness: TC, vertex form.  Same argv (input_graph verify)
// and stdout contract as the reference's include/tc_vertex_cuda.h main(): median of 3
// timed calls, one "GPU Throughput: <GTEPS> gigaedges/s" line, the count
// checked against a host merge count when asked.
#include "common.hpp"

typedef long long basic_t;

int main(int argc, char* argv[])
{
  printf("Triangle counting vertex-centric HIP (%s)\n", __FILE__);
  if (argc != 3) {
    printf("USAGE: %s input_graph verify\n\n", argv[0]);
    exit(-1);
  }
  ECLgraph g = readECLgraph(argv[1]);
  printf("graph: %s\n", argv[1]);
  printf("nodes: %d\n", g.nodes);
  printf("edges: %d\n", g.edges);
  printf("avg degree: %.2f\n", 1.0 * g.edges / g.nodes);
  int mdeg = 0;
  for (int v = 0; v < g.nodes; v++) mdeg = std::max(mdeg, g.nindex[v + 1] - g.nindex[v]);
  printf("max degree: %d\n\n", mdeg);
  harness::require_sorted(g);
  GPUinfo(0);
  basic_t count = 0;
  ECLgraph d_g = harness::upload(g, false);
  const double med = harness::timed_runs(3, [&] { return GPUtc_vertex(count, d_g.nodes, d_g.nindex, d_g.nlist); });
  printf("GPU runtime: %.6f s\n", med);
  printf("GPU Throughput: %.6f gigaedges/s\n", 0.000000001 * g.edges / med);
  const int runveri = harness::verify_flag(argv[2]);
  if (runveri) {
    const basic_t h_count = harness::host_triangles(g);
    if (h_count != count) printf("ERROR: host %lld device %lld", h_count, count);
    else printf("the pattern occurs %lld times\n\n", count);
  } else {
    printf("turn off verification\n\n");
  }
  harness::release(d_g);
  freeECLgraph(g);
  return 0;
}
