/*
 * indigo_graph.h - host-side CSR construction, synthetic graph generators and
 * .egr file I/O (C ABI).  Host code only; no GPU is touched.
 *
 * Graphs are returned in an indigo_host_graph whose layout equals the
 * reference's struct ECLgraph (include/ECLgraph.h:44-50): int nodes, int edges,
 * int* nindex, int* nlist, int* eweight.  The arrays are malloc'd by the
 * library and released with indigo_graph_release().
 *
 * Every generated graph is simple and undirected in CSR form: self loops
 * dropped, each undirected edge stored in both directions, duplicates merged
 * and each adjacency list sorted ascending - the form the reference's TC
 * harness requires (include/tc_vertex_cuda.h:157-165).  Weights, when asked
 * for (wmax >= wmin >= 1), are uniform integers in [wmin, wmax] drawn per
 * undirected edge from (seed, min(u,v), max(u,v)), so both directions carry
 * the same weight.  All generators are deterministic in their seed and
 * independent of the thread count.
 */
#ifndef INDIGO_GRAPH_H
#define INDIGO_GRAPH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct indigo_host_graph {
  int nodes;
  int edges;
  int* nindex;  /* nodes + 1 */
  int* nlist;   /* edges */
  int* eweight; /* edges, or NULL */
} indigo_host_graph;

/* Build a simple undirected CSR from npairs (src[i], dst[i]) pairs.
 * symmetrize = 1 adds the reverse of every pair.  wmax < wmin or wmin < 1
 * means no weights.  threads <= 0 picks the hardware concurrency (max 16). */
int indigo_build_csr(int nodes, long long npairs, const int* src, const int* dst, int symmetrize,
                     unsigned long long wseed, int wmin, int wmax, int threads, indigo_host_graph* out);

/* rows x cols 4-neighbour grid; vertex (r, c) has id r * cols + c. */
int indigo_gen_grid(int rows, int cols, unsigned long long seed, int wmin, int wmax, indigo_host_graph* out);

/* R-MAT with 2^scale vertices and edge_factor * 2^scale samples, quadrant
 * probabilities a, b, c (d = 1 - a - b - c), no vertex permutation (vertex 0
 * is the densest row). */
int indigo_gen_rmat(int scale, int edge_factor, double a, double b, double c, unsigned long long seed,
                    int wmin, int wmax, int threads, indigo_host_graph* out);

/* Erdos-Renyi G(n, m) style: `samples` uniform vertex pairs. */
int indigo_gen_er(int nodes, long long samples, unsigned long long seed, int wmin, int wmax, int threads,
                  indigo_host_graph* out);

/* Road-like grid: each 4-neighbour edge kept with probability keep_p, plus,
 * for each vertex with probability diag_p, one edge to a uniformly chosen
 * in-bounds diagonal neighbour. */
int indigo_gen_road(int rows, int cols, double keep_p, double diag_p, unsigned long long seed, int wmin,
                    int wmax, int threads, indigo_host_graph* out);

/* Deep copy (arrays malloc'd) and release. */
int indigo_graph_copy(const indigo_host_graph* src, indigo_host_graph* out);
int indigo_graph_release(indigo_host_graph* g);

/* Degree-ordered relabelling (no counterpart in the reference: a graph
 * preparation step for callers that query one graph many times).  New ids go
 * by degree, largest first, ties in id order; perm[v] (nodes ints, caller's
 * memory) receives the new id of vertex v.  out holds the same graph under
 * the new ids, each list sorted by new id, weights moved with their entries.
 * Hubs then sit at the smallest ids, so the per-vertex data that many lists
 * gather (distances, PageRank contributions) shares cache lines.  Results of
 * BFS, SSSP and PageRank on out, read through perm (indigo_unpermute), equal
 * those on src; CC labels and MIS sets depend on the ids and are not
 * carried over. */
int indigo_graph_reorder(const indigo_host_graph* src, int threads, indigo_host_graph* out, int* perm);

/* .egr I/O in the ECLgraph format (reference include/ECLgraph.h:52-96);
 * unlike the header-only loader these return status codes instead of exiting. */
int indigo_read_egr(const char* path, indigo_host_graph* out);
int indigo_write_egr(const char* path, const indigo_host_graph* g);

/* A read-only memory map of an .egr file: the arrays point into the
 * mapping, so nothing is read until it is touched (the reference's
 * readECLgraph, include/ECLgraph.h:52-77, freads the whole file into
 * malloc'd arrays).  eweight is NULL when the file carries no weights. */
typedef struct indigo_egr_map {
  int nodes;
  int edges;
  const int* nindex;
  const int* nlist;
  const int* eweight;
  void* base;    /* the mapping */
  size_t length; /* its length in bytes */
} indigo_egr_map;

int indigo_egr_open(const char* path, indigo_egr_map* out);
int indigo_egr_close(indigo_egr_map* m);

/* Checks the invariants above (sorted, no self loops, symmetric, weights
 * symmetric).  Returns 0 if they hold, else a positive code naming the first
 * violation (1 index, 2 range, 3 unsorted/duplicate, 4 self loop,
 * 5 asymmetric, 6 asymmetric weight). */
int indigo_graph_check(const indigo_host_graph* g);

#ifdef __cplusplus
}
#endif

#endif /* INDIGO_GRAPH_H */
