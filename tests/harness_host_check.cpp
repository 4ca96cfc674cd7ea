// Host parts of the drop-in harnesses (parallelization-artifact-sc23_amd/
// harness/common.hpp) run on a graph file, without a GPU: the weight
// synthesis for graphs stored without weights, the edge-source array and
// the median of the timed runs.  Built by tests/test_harness_host.py, which
// checks the outputs against its own restatement of the reference's rules.
//
// Usage: harness_host_check <graph.egr> <out.bin>
//   out.bin: int32 weights[edges], int32 sp[edges], then float64 medians of
//   {3, 1, 2}, {4, 1, 3, 2} and {7}.
#include <cstdio>
#include <vector>

#include "common.hpp"

int main(int argc, char** argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <graph.egr> <out.bin>\n", argv[0]);
    return 2;
  }
  ECLgraph g = readECLgraph(argv[1]);
  if (g.eweight != nullptr) {
    std::fprintf(stderr, "expected a graph without weights\n");
    return 2;
  }
  harness::synthesize_weights(g);
  const std::vector<int> sp = harness::edge_sources(g);
  const double med[3] = {harness::median({3.0, 1.0, 2.0}), harness::median({4.0, 1.0, 3.0, 2.0}),
                         harness::median({7.0})};
  FILE* f = std::fopen(argv[2], "wb");
  if (f == nullptr) return 1;
  const size_t e = static_cast<size_t>(g.edges);
  bool ok = std::fwrite(g.eweight, sizeof(int), e, f) == e;
  ok = ok && std::fwrite(sp.data(), sizeof(int), e, f) == e;
  ok = ok && std::fwrite(med, sizeof(double), 3, f) == 3;
  ok = (std::fclose(f) == 0) && ok;
  freeECLgraph(g);
  return ok ? 0 : 1;
}
