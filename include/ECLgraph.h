/*
 * ECLgraph CSR container and binary .egr reader/writer.
 *
 * Same struct layout, function names and on-disk format as the Indigo2 suite's
 * loader (include/ECLgraph.h:44-106 of the reference), so that drop-in harnesses
 * and tools that expect that API compile unchanged.  The implementation here is
 * independent.
 *
 * File format (little-endian int32 throughout):
 *   nodes, edges, nindex[nodes + 1], nlist[edges], [eweight[edges]]
 * The weight block is optional: a file that ends after nlist loads with
 * eweight == NULL (reference ECLgraph.h:70-76); a partial weight block is an error.
 *
 * Error behaviour mirrors the reference: any I/O or allocation failure prints
 * "ERROR: ..." to stderr and calls exit(-1).
 */
#ifndef ECL_GRAPH
#define ECL_GRAPH

#include <cstdio>
#include <cstdlib>

struct ECLgraph {
  int nodes;
  int edges;
  int* nindex;
  int* nlist;
  int* eweight;
};

namespace ecl_detail {

[[noreturn]] inline void fail(const char* what, const char* arg = nullptr)
{
  if (arg) fprintf(stderr, "ERROR: %s %s\n\n", what, arg);
  else fprintf(stderr, "ERROR: %s\n\n", what);
  exit(-1);
}

template <typename T>
inline T* alloc_array(long long count)
{
  // malloc(0) may legally return NULL; always ask for at least one element so
  // an edgeless graph still gets valid (empty) arrays.
  T* p = static_cast<T*>(malloc(sizeof(T) * static_cast<size_t>(count > 0 ? count : 1)));
  if (p == nullptr) fail("memory allocation failed");
  return p;
}

}  // namespace ecl_detail

inline ECLgraph readECLgraph(const char* const fname)
{
  using namespace ecl_detail;
  FILE* f = fopen(fname, "rb");
  if (f == nullptr) fail("could not open file", fname);

  ECLgraph g;
  if (fread(&g.nodes, sizeof(int), 1, f) != 1) fail("failed to read nodes");
  if (fread(&g.edges, sizeof(int), 1, f) != 1) fail("failed to read edges");
  printf("input graph: %d nodes and %d edges\n", g.nodes, g.edges);
  if (g.nodes < 1 || g.edges < 0) fail("node or edge count too low");

  g.nindex = alloc_array<int>(static_cast<long long>(g.nodes) + 1);
  g.nlist = alloc_array<int>(g.edges);
  g.eweight = alloc_array<int>(g.edges);

  if (fread(g.nindex, sizeof(int), static_cast<size_t>(g.nodes) + 1, f) != static_cast<size_t>(g.nodes) + 1)
    fail("failed to read neighbor index list");
  if (fread(g.nlist, sizeof(int), static_cast<size_t>(g.edges), f) != static_cast<size_t>(g.edges))
    fail("failed to read neighbor list");
  const size_t got = fread(g.eweight, sizeof(int), static_cast<size_t>(g.edges), f);
  if (got == 0) {
    free(g.eweight);
    g.eweight = nullptr;
  } else if (got != static_cast<size_t>(g.edges)) {
    fail("failed to read edge weights");
  }
  fclose(f);
  return g;
}

inline void writeECLgraph(const ECLgraph g, const char* const fname)
{
  using namespace ecl_detail;
  if (g.nodes < 1 || g.edges < 0) fail("node or edge count too low");
  FILE* f = fopen(fname, "wb");
  if (f == nullptr) fail("could not open file", fname);
  if (fwrite(&g.nodes, sizeof(int), 1, f) != 1) fail("failed to write nodes");
  if (fwrite(&g.edges, sizeof(int), 1, f) != 1) fail("failed to write edges");
  if (fwrite(g.nindex, sizeof(int), static_cast<size_t>(g.nodes) + 1, f) != static_cast<size_t>(g.nodes) + 1)
    fail("failed to write neighbor index list");
  if (fwrite(g.nlist, sizeof(int), static_cast<size_t>(g.edges), f) != static_cast<size_t>(g.edges))
    fail("failed to write neighbor list");
  if (g.eweight != nullptr &&
      fwrite(g.eweight, sizeof(int), static_cast<size_t>(g.edges), f) != static_cast<size_t>(g.edges))
    fail("failed to write edge weights");
  fclose(f);
}

inline void freeECLgraph(ECLgraph& g)
{
  free(g.nindex);
  free(g.nlist);
  free(g.eweight);
  g.nindex = nullptr;
  g.nlist = nullptr;
  g.eweight = nullptr;
}

#endif  // ECL_GRAPH
