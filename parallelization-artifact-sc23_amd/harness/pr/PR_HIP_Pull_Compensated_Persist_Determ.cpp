// Drop-in harness: PageRank.  Same argv (input_graph) and stdout contract as
// the reference's include/pr_cuda.h main(): PR_GPU prints the iteration
// count and the Throughput line; the result is checked by one more Jacobi
// step on the host (total change below 1e-4: "All good.").
#include "common.hpp"

typedef float score_type;

int main(int argc, char* argv[])
{
  printf("PageRank HIP (%s)\n\n", __FILE__);
  if (argc != 2) {
    printf("USAGE: %s input_graph\n\n", argv[0]);
    exit(-1);
  }
  ECLgraph g = readECLgraph(argv[1]);
  printf("graph: %s\n", argv[1]);
  printf("nodes: %d\n", g.nodes);
  printf("edges: %d\n", g.edges);
  printf("avg degree: %.2f\n", 1.0 * g.edges / g.nodes);
  GPUinfo(0);
  std::vector<int> degree(static_cast<size_t>(g.nodes));
  for (int v = 0; v < g.nodes; v++) degree[v] = g.nindex[v + 1] - g.nindex[v];
  std::vector<score_type> scores(static_cast<size_t>(g.nodes), 1.0f / static_cast<score_type>(g.nodes));
  PR_GPU(g, scores.data(), degree.data());
  const score_type damp = 0.85f;
  const score_type base = (1.0f - damp) / static_cast<score_type>(g.nodes);
  std::vector<score_type> incoming(static_cast<size_t>(g.nodes), 0.0f);
  for (int u = 0; u < g.nodes; u++) {
    if (degree[u] == 0) continue;
    const score_type contrib = scores[u] / degree[u];
    for (int j = g.nindex[u]; j < g.nindex[u + 1]; j++) incoming[g.nlist[j]] += contrib;
  }
  double error = 0.0;
  for (int v = 0; v < g.nodes; v++) error += std::fabs(base + damp * incoming[v] - scores[v]);
  if (error < 0.0001) printf("All good.\n");
  else printf("Total Error: %f\n", error);
  freeECLgraph(g);
  return 0;
}
