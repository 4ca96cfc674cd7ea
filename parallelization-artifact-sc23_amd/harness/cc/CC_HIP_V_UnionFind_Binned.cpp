// Drop-in harness: CC, vertex form.  Same argv (input_file_name
// verify) and stdout contract as the reference's include/cc_vertex_cuda.h main():
// median of 3 timed calls, one "Throughput: <GTEPS> gigaedges/s" line.
#include <set>

#include "common.hpp"

typedef int basic_t;

int main(int argc, char* argv[])
{
  printf("cc vertex-centric HIP (%s)\n", __FILE__);
  if (argc != 3) {
    fprintf(stderr, "USAGE: %s input_file_name verify\n", argv[0]);
    exit(-1);
  }
  ECLgraph g = readECLgraph(argv[1]);
  printf("input: %s\n", argv[1]);
  printf("nodes: %d\n", g.nodes);
  printf("edges: %d\n", g.edges);
  const int runveri = harness::verify_flag(argv[2]);
  GPUinfo(0);
  std::vector<basic_t> label(static_cast<size_t>(g.nodes));
  ECLgraph d_g = harness::upload(g, false);
  const double med = harness::timed_runs(3, [&] { return GPUcc_vertex(d_g, label.data()); });
  harness::report(med, g.edges);
  const std::set<int> components(label.begin(), label.end());
  printf("number of connected components: %zu\n", components.size());
  if (runveri) harness::verify_components(g, label.data());
  else printf("turn off verification\n\n");
  harness::release(d_g);
  freeECLgraph(g);
  return 0;
}
