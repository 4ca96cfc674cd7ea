// Shared parts of the drop-in harness mains (one executable per entry form,
// same argv and stdout contract as the reference harnesses in
// include/{alg}_cuda.h): device upload of an ECLgraph, the timed runs and
// their median, the Throughput line, the reference's weight synthesis for
// graphs stored without weights, the edge-source array, and verification of
// each output by the properties that define it.  Nothing here computes a
// result the GPU is then trusted for: verification only checks.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#include "ECLgraph.h"
#include "indigo_compat.hpp"

namespace harness {

inline int* device_ints(const int* h, size_t n)
{
  int* d = nullptr;
  indigo_compat::check(indigo_malloc(reinterpret_cast<void**>(&d), n * sizeof(int)), "device allocation");
  if (n > 0) indigo_compat::check(indigo_memcpy_h2d(d, h, n * sizeof(int)), "copy to device");
  return d;
}

// The harness's d_g: host sizes, device arrays (bfs_vertex_cuda.h:140-144).
inline ECLgraph upload(const ECLgraph& g, bool weights)
{
  ECLgraph d = g;
  d.nindex = device_ints(g.nindex, static_cast<size_t>(g.nodes) + 1);
  d.nlist = device_ints(g.nlist, static_cast<size_t>(g.edges));
  d.eweight = weights ? device_ints(g.eweight, static_cast<size_t>(g.edges)) : nullptr;
  return d;
}

// Frees the device graph and releases the library's default context (its
// stream, workspace and pinned buffers) while the HIP runtime is still up,
// before main returns.
inline void release(ECLgraph& d)
{
  indigo_free(d.nindex);
  indigo_free(d.nlist);
  if (d.eweight != nullptr) indigo_free(d.eweight);
  indigo_shutdown();
}

// Median as the reference harness takes it (bfs_vertex_cuda.h:111-118).
inline double median(std::vector<double> v)
{
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  return n % 2 == 0 ? (v[(n - 1) / 2] + v[n / 2]) / 2.0 : v[n / 2];
}

// `runs` timed calls; a call of 30 s or more ends the series and is
// reported alone (bfs_vertex_cuda.h:150-157).
inline double timed_runs(int runs, const std::function<double()>& call)
{
  std::vector<double> t;
  for (int i = 0; i < runs; i++) {
    t.push_back(call());
    if (t.back() >= 30.0) return t.back();
  }
  return median(t);
}

inline void report(double seconds, int edges)
{
  printf("runtime: %.6fs\n", seconds);
  printf("Throughput: %.6f gigaedges/s\n", 0.000000001 * edges / seconds);
}

inline int verify_flag(const char* arg)
{
  const int v = atoi(arg);
  if (v != 0 && v != 1) printf("has to be 0 (turn off) or 1 (turn on) verification");
  return v;
}

// Weights for a graph stored without them, by the reference harness's rule
// (sssp_vertex_cuda.h:127-137): w = |1 + (i * nei) mod nodes| with the
// product wrapping in 32 bits.
inline void synthesize_weights(ECLgraph& g)
{
  if (g.eweight != nullptr) return;
  printf("Generating weights.\n");
  g.eweight = static_cast<int*>(malloc(sizeof(int) * static_cast<size_t>(g.edges > 0 ? g.edges : 1)));
  for (int i = 0; i < g.nodes; i++) {
    for (int j = g.nindex[i]; j < g.nindex[i + 1]; j++) {
      const int prod = static_cast<int>(static_cast<unsigned>(i) * static_cast<unsigned>(g.nlist[j]));
      int w = 1 + prod % g.nodes;
      g.eweight[j] = w < 0 ? -w : w;
    }
  }
}

// sp[j] = v for j in [nindex[v], nindex[v+1]) (bfs_edge_cuda.h:140-146).
inline std::vector<int> edge_sources(const ECLgraph& g)
{
  std::vector<int> sp(static_cast<size_t>(g.edges));
  for (int v = 0; v < g.nodes; v++)
    for (int j = g.nindex[v]; j < g.nindex[v + 1]; j++) sp[static_cast<size_t>(j)] = v;
  return sp;
}

inline void fail(const char* msg, int v)
{
  fprintf(stderr, "ERROR: verification failed for node %d: %s\n", v, msg);
  exit(-1);
}

// Shortest-path distances on a symmetric graph (unit weights when
// `weighted` is false) are exactly the values with dist[src] = 0, no edge
// violating dist[u] <= dist[v] + w, and a tight edge into every other
// reached vertex; unreached vertices (INT_MAX) have no reached neighbour.
inline void verify_paths(const ECLgraph& g, int src, const int* dist, bool weighted)
{
  if (dist[src] != 0) fail("source distance is not 0", src);
  for (int v = 0; v < g.nodes; v++) {
    const int dv = dist[v];
    bool tight = v == src;
    for (int j = g.nindex[v]; j < g.nindex[v + 1]; j++) {
      const int u = g.nlist[j];
      const long long w = weighted ? g.eweight[j] : 1;
      const int du = dist[u];
      if ((dv == INT_MAX) != (du == INT_MAX)) fail("reached vertex next to an unreached one", v);
      if (dv == INT_MAX) continue;
      if (static_cast<long long>(du) > dv + w) fail("edge violates the triangle inequality", u);
      if (static_cast<long long>(du) + w == dv) tight = true;
    }
    if (dv != INT_MAX && (dv < 0 || !tight)) fail("distance not attained by any neighbour", v);
  }
  printf("verification passed\n\n");
}

// Connected-component labels must equal the smallest id of each component:
// labelling components by a host BFS started from each unlabelled vertex in
// id order gives exactly that.
inline void verify_components(const ECLgraph& g, const int* label)
{
  std::vector<int> want(static_cast<size_t>(g.nodes), -1), queue;
  for (int r = 0; r < g.nodes; r++) {
    if (want[r] >= 0) continue;
    want[r] = r;
    queue.assign(1, r);
    for (size_t k = 0; k < queue.size(); k++) {
      const int v = queue[k];
      for (int j = g.nindex[v]; j < g.nindex[v + 1]; j++)
        if (want[g.nlist[j]] < 0) {
          want[g.nlist[j]] = r;
          queue.push_back(g.nlist[j]);
        }
    }
  }
  for (int v = 0; v < g.nodes; v++)
    if (label[v] != want[v]) fail("label is not the smallest id of the component", v);
  printf("verification passed\n\n");
}

// MIS status (1 included, 2 excluded): all decided, independent, maximal
// (the reference's checks, mis_vertex_cuda.h:80-97).
inline void verify_mis(const ECLgraph& g, const int* status)
{
  for (int v = 0; v < g.nodes; v++) {
    if (status[v] != 1 && status[v] != 2) {
      fprintf(stderr, "ERROR: found undecided node\n");
      exit(-1);
    }
    bool covered = status[v] == 1;
    for (int j = g.nindex[v]; j < g.nindex[v + 1]; j++) {
      if (status[v] == 1 && status[g.nlist[j]] == 1) {
        fprintf(stderr, "ERROR: found adjacent nodes in MIS\n");
        exit(-1);
      }
      covered |= status[g.nlist[j]] == 1;
    }
    if (!covered) {
      fprintf(stderr, "ERROR: set is not maximal\n");
      exit(-1);
    }
  }
  printf("verification passed\n\n");
}

// Triangle count on the host by sorted merges over id-ordered lists.
inline long long host_triangles(const ECLgraph& g)
{
  long long count = 0;
  for (int v = 0; v < g.nodes; v++) {
    const int* a = g.nlist + g.nindex[v];
    const int* ae = g.nlist + g.nindex[v + 1];
    const int* ahi = std::upper_bound(a, ae, v);  // neighbours above v
    for (const int* p = ahi; p < ae; p++) {
      const int u = *p;
      const int* b = std::upper_bound(g.nlist + g.nindex[u], g.nlist + g.nindex[u + 1], u);
      const int* be = g.nlist + g.nindex[u + 1];
      const int* x = p + 1;
      while (x < ae && b < be) {
        if (*x < *b) x++;
        else if (*b < *x) b++;
        else {
          count++;
          x++;
          b++;
        }
      }
    }
  }
  return count;
}

// TC input must have sorted lists without self loops (tc_vertex_cuda.h:157-164).
inline void require_sorted(const ECLgraph& g)
{
  for (int v = 0; v < g.nodes; v++)
    for (int i = g.nindex[v]; i < g.nindex[v + 1]; i++)
      if (g.nlist[i] == v || (i > g.nindex[v] && g.nlist[i - 1] >= g.nlist[i])) {
        printf("ERROR: adjacency list not sorted or contains self edge\n");
        exit(-1);
      }
}

}  // namespace harness
