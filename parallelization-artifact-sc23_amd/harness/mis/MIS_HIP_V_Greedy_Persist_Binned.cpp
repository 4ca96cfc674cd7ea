// Drop-in harness: MIS, vertex form.  Same argv (input_file_name
// verify) and stdout contract as the reference's include/mis_vertex_cuda.h main():
// median of 3 timed calls, each verified when asked, one "Throughput" line.
// Priorities are unsigned (the reference's C++ order); building with
// -DINDIGO_MIS_SIGNED selects the CUDA variants' signed order.
#include "common.hpp"

#ifdef INDIGO_MIS_SIGNED
typedef int data_type;
#else
typedef unsigned data_type;
#endif

int main(int argc, char* argv[])
{
  printf("mis vertex-based HIP (%s)\n", __FILE__);
  if (argc != 3) {
    fprintf(stderr, "USAGE: %s input_file_name verify\n", argv[0]);
    exit(-1);
  }
  ECLgraph g = readECLgraph(argv[1]);
  harness::synthesize_weights(g);  // as the reference harness does; MIS itself does not read them
  printf("input: %s\n", argv[1]);
  printf("nodes: %d\n", g.nodes);
  printf("edges: %d\n", g.edges);
  const int runveri = harness::verify_flag(argv[2]);
  GPUinfo(0);
  std::vector<data_type> priority(static_cast<size_t>(g.nodes));
  std::vector<int> status(static_cast<size_t>(g.nodes));
  ECLgraph d_g = harness::upload(g, true);
  const double med = harness::timed_runs(3, [&] {
    const double s = GPUmis_vertex(d_g, priority.data(), status.data());
    if (runveri) harness::verify_mis(g, status.data());
    else printf("turn off verification\n\n");
    return s;
  });
  harness::report(med, g.edges);
  int size = 0;
  for (int v = 0; v < g.nodes; v++) size += status[v] == 1;
  printf("elements in set: %d (%.1f%%)\n", size, 100.0 * size / g.nodes);
  harness::release(d_g);
  freeECLgraph(g);
  return 0;
}
