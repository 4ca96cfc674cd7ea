/*
 * indigo_hip.h - C ABI of the MI355X (gfx950) implementation of Indigo2's six
 * graph kernels: BFS, SSSP, CC, MIS, PageRank and triangle counting over the
 * ECLgraph CSR (nindex / nlist / eweight, all int32).
 *
 * Every function returns an int status (INDIGO_OK == 0, negative on failure);
 * indigo_last_error() returns a per-thread description of the last failure.
 * Pointers are plain host or device pointers as stated per argument ("d_" =
 * device memory, "h_" = host memory); no function keeps a pointer after it
 * returns.
 *
 * Two layers:
 *   1. Drop-in entries (indigo_bfs_vertex ... indigo_tc_edge): one per entry
 *      function of the reference harness headers, with the same pointer
 *      ownership, the same outputs and the same timing convention (wall time
 *      around the iteration loop, allocation and H2D of the graph excluded).
 *      The C++ harnesses wrap them in the reference's static signatures.
 *   2. Resident entries (indigo_ctx_*, indigo_bfs ... indigo_tc): the graph
 *      and the outputs stay in device memory, work is issued on the context's
 *      stream with cached workspace, and the kernel time of the last call is
 *      measured with HIP events on that stream.  Benchmarks, the batched
 *      engine and the serving loop use this layer.
 *
 * Graph contents are trusted, as the reference trusts them: nindex[0] = 0,
 * nindex non-decreasing up to nindex[nodes] = edges, every nlist entry in
 * [0, nodes), and symmetric lists for CC, MIS and TC.  Sizes, sources and
 * null pointers are checked; list contents are not, since that would cost a
 * pass over the edges per call (TC alone reports unsorted lists and self
 * loops, with INDIGO_ERR_UNSORTED, as the reference's TC does).
 * indigo_graph_check (indigo_graph.h) verifies a host graph beforehand.
 */
#ifndef INDIGO_HIP_H
#define INDIGO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define INDIGO_OK 0
#define INDIGO_ERR_INVALID_ARG (-1)   /* null pointer, bad size, source out of range */
#define INDIGO_ERR_OUT_OF_MEMORY (-2) /* device or host allocation failed */
#define INDIGO_ERR_DEVICE (-3)        /* HIP runtime error (launch, copy, sticky fault) */
#define INDIGO_ERR_UNSORTED (-4)      /* TC input has unsorted lists or self loops */
#define INDIGO_ERR_TIMEOUT (-5)       /* a bounded device-side wait expired; results invalid */
#define INDIGO_ERR_CAPACITY (-6)      /* request exceeds a fixed capacity (batch, admission) */
#define INDIGO_ERR_NO_DEVICE (-7)     /* no usable GPU */

/* Description of the last error on the calling thread ("" if none). */
const char* indigo_last_error(void);

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int indigo_version(void);

/* Device properties.  Replaces GPUinfo (include/bfs_vertex_cuda.h:63-73):
 * *compute_units x *max_threads_per_cu is the persistent-grid bound the
 * reference derives from SMs x maxThreadsPerSM.  name may be NULL. */
int indigo_device_info(int device, int* compute_units, int* max_threads_per_cu, char* name, int name_len);

/* Device memory plumbing for harnesses (cudaMalloc / cudaMemcpy / cudaFree
 * call sites in every include/{alg}_cuda.h main(), e.g. bfs_vertex_cuda.h:140-144). */
int indigo_malloc(void** d_ptr, size_t bytes);
int indigo_free(void* d_ptr);
int indigo_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes);
int indigo_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes);

/* ---- drop-in entries ---------------------------------------------------- */
/* Each writes the elapsed seconds of its iteration loop to *seconds (may be
 * NULL).  Unreachable vertices get INT_MAX (reference bfs_vertex_cuda.h:46). */

/* GPUbfs_vertex, include/bfs_vertex_cuda.h:61: hop distance from src.
 * d_nindex/d_nlist device, h_dist host int[nodes]. */
int indigo_bfs_vertex(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                      int* h_dist, double* seconds);

/* GPUbfs_edge, include/bfs_edge_cuda.h:61: same output, computed by the
 * edge-centric, topology-driven kernel (one lane per CSR entry): h_sp is the
 * host edge-source array sp[e] (bfs_edge_cuda.h:140-146), copied to the
 * device before the timer starts; its contents are trusted like the CSR's. */
int indigo_bfs_edge(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                    const int* h_sp, int* h_dist, double* seconds);

/* GPUsssp_vertex, include/sssp_vertex_cuda.h:60: int32 shortest-path
 * distances with the device weights d_eweight[edges]. */
int indigo_sssp_vertex(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                       const int* d_eweight, int* h_dist, double* seconds);

/* GPUsssp_edge, include/sssp_edge_cuda.h:60: edge-centric Bellman-Ford sweeps
 * over (h_sp[e], d_nlist[e], d_eweight[e]); h_sp host, as in the reference. */
int indigo_sssp_edge(int src, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                     const int* d_eweight, const int* h_sp, int* h_dist, double* seconds);

/* GPUcc_vertex, include/cc_vertex_cuda.h:57: label[v] = smallest vertex id
 * in v's connected component. */
int indigo_cc_vertex(int nodes, int edges, const int* d_nindex, const int* d_nlist,
                     int* h_label, double* seconds);

/* GPUcc_edge, include/cc_edge_cuda.h:57: one union-find linking sweep over
 * the entries (h_sp[e], d_nlist[e]) with h_sp[e] > d_nlist[e] (h_sp host). */
int indigo_cc_edge(int nodes, int edges, const int* d_nindex, const int* d_nlist,
                   const int* h_sp, int* h_label, double* seconds);

/* GPUmis_vertex, include/mis_vertex_cuda.h:54: status[v] in {1 included,
 * 2 excluded} of the greedy maximal independent set in descending
 * (priority, id) order, priority[v] = hash(v + 712313887)
 * (mis_vertex_cuda.h:174-179).  signed_priority = 0 compares priorities as
 * unsigned int (the suite's C++ path, the default), 1 as signed int (the
 * suite's CUDA path).  h_priority (host unsigned[nodes]) receives the
 * priorities if non-NULL. */
int indigo_mis_vertex(int nodes, int edges, const int* d_nindex, const int* d_nlist,
                      int signed_priority, unsigned* h_priority, int* h_status, double* seconds);

/* GPUmis_edge, include/mis_edge_cuda.h:54: rounds of edge-centric sweeps
 * (candidates withdrawn by larger-key neighbours, neighbours of members
 * excluded) over d_sp (device, as in the reference). */
int indigo_mis_edge(int nodes, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                    int signed_priority, unsigned* h_priority, int* h_status, double* seconds);

/* PR_GPU, include/pr_cuda.h:48: the graph is on the HOST and the callee
 * copies it.  h_scores (host float[nodes]) holds the initial scores on entry
 * (the harness fills 1/nodes) and the result on exit; h_degree host int[nodes].
 * Damping 0.85; stops when the L1 change of an iteration is <= 1e-4 or after
 * 100 iterations (pr_cuda.h:42-45).  *iterations (may be NULL) gets the count. */
int indigo_pr(int nodes, int edges, const int* h_nindex, const int* h_nlist, const int* h_degree,
              float* h_scores, int* iterations, double* seconds);

/* GPUtc_vertex, include/tc_vertex_cuda.h:42: number of triangles of a graph
 * with sorted, self-loop-free adjacency lists, counted in 64 bits. */
int indigo_tc_vertex(long long* count, int nodes, const int* d_nindex, const int* d_nlist, double* seconds);

/* GPUtc_edge, include/tc_edge_cuda.h:42 (d_sp device, sp sorted as built by
 * the harness, tc_edge_cuda.h:159-165). */
int indigo_tc_edge(long long* count, int edges, const int* d_nindex, const int* d_nlist, const int* d_sp,
                   double* seconds);

/* ---- resident layer ---------------------------------------------------- */
typedef struct indigo_ctx indigo_ctx;

/* A context owns one HIP stream (default priority), a workspace cache and a
 * pair of timing events.  Contexts are independent; one context must not be
 * used by two host threads at once. */
int indigo_ctx_create(int device, indigo_ctx** out);
int indigo_ctx_destroy(indigo_ctx* ctx);
/* Releases the process-wide context of the drop-in entries (its stream,
 * workspace and pinned buffers) and waits for every device to go idle.  Call
 * it before the process exits, while the HIP runtime is still up: the
 * harness mains do at the end of main, the Python binding from an atexit
 * hook.  A later drop-in call creates the context again.  Contexts made
 * with indigo_ctx_create are the caller's to destroy.  No counterpart in the
 * reference (its harnesses leave teardown to process exit). */
int indigo_shutdown(void);
/* Milliseconds between the first and last kernel of the last resident call,
 * measured with HIP events on the context stream. */
int indigo_ctx_last_ms(indigo_ctx* ctx, float* ms);
/* Number of kernel launches and iterations (levels, rounds, sweeps) of the
 * last resident call. */
int indigo_ctx_last_stats(indigo_ctx* ctx, int* launches, int* iterations);
/* Duration in milliseconds (HIP events on the context stream, around that one
 * launch) and name of the main kernel of the last resident call: bfs_kernel,
 * sssp_kernel, cc_link, mis_kernel, pr_kernel, or tc_count_group<512, 8192>.
 * name (name_len bytes, may be NULL) receives "" and ms the whole call when
 * the call marked no kernel. */
int indigo_ctx_last_kernel(indigo_ctx* ctx, float* ms, char* name, int name_len);
/* Device-to-device copy of `bytes` on the context stream (hipMemcpyAsync),
 * timed with the context's events: the copy-bandwidth reference beside the
 * kernels' roofline figures.  No counterpart in the reference. */
int indigo_ctx_copy_ms(indigo_ctx* ctx, void* d_dst, const void* d_src, size_t bytes, float* ms);
/* Workgroups of the last persistent launch and the per-CU
 * residency it was sized for. */
int indigo_ctx_last_grid(indigo_ctx* ctx, int* grid, int* per_cu);
/* Per-level trace of the last BFS or SSSP call (first 256 levels): start of
 * each level in microseconds since the kernel began, frontier size, its
 * degree sum, and the mode (0 top-down, 1 bottom-up, 2 one-workgroup phase,
 * 3 SSSP far-pile split).  After an MIS call, one entry per round: its start,
 * the undecided vertices with short lists and (as degree_sum) those with
 * long lists.  Any output pointer may be NULL. */
int indigo_ctx_last_trace(indigo_ctx* ctx, int max, float* us, unsigned* frontier, double* degree_sum, int* mode,
                          int* n);
/* Release the cached workspace (it regrows on the next call). */
int indigo_ctx_trim(indigo_ctx* ctx);
/* Optional tuning knob: block size / variant selection hints (0 = auto). */
int indigo_ctx_set_hint(indigo_ctx* ctx, int key, int value);
#define INDIGO_HINT_BFS_DIRECTION 1 /* 0 auto, 1 top-down only, 2 allow bottom-up */
#define INDIGO_HINT_SSSP_DELTA 2    /* 0 auto, else near-far split increment */
#define INDIGO_HINT_SMALL_FRONTIER 3 /* 0 auto, 1 off: never use the one-workgroup mode (BFS, SSSP) */
#define INDIGO_HINT_BUDGET_MS 4      /* bound on device-side waits in ms (0: 60000) */
#define INDIGO_HINT_BLOCKS_PER_CU 5  /* cap on co-resident workgroups per CU of persistent kernels (0: occupancy) */
#define INDIGO_HINT_TC_LDS_LIST 6    /* entries of an oriented list TC hashes in LDS at once (0: 2048 or 8192 by
                                        list length); longer lists are hashed that many at a time */
#define INDIGO_HINT_RESIDENCY_MS 7   /* how long a persistent grid may take to assemble (0: 100 ms); longer when
                                        the GPU is shared with other work */
#define INDIGO_HINT_MAX_WORKGROUPS 8 /* cap on the grids of the context, in workgroups of 512 threads (other
                                        sizes: as many threads); 0: the whole GPU.  For contexts that share
                                        the GPU (each serving slot holds its share) */
#define INDIGO_HINT_VARIANT 9        /* INDIGO_VARIANT_*: which form indigo_bfs / _sssp / _cc / _mis run */
#define INDIGO_HINT_BLOCK_SIZE 10    /* workgroup size of the edge-centric kernels: 256, 512 (0: default) or 1024 */
#define INDIGO_HINT_GRAPH 11        /* 1: CC and TC run their kernel sequence as a HIP graph, captured on the
                                        first call with the same graph, output and workspace and replayed after
                                        (one launch instead of 5-10); such calls report no separate main kernel */
#define INDIGO_HINT_TC_HUBS 12       /* cap on TC's hub vertices, whose out-lists hold 16-bit keys (0: 65535) */
#define INDIGO_VARIANT_VERTEX 0      /* vertex-centric, data-driven (degree-binned worklists); the default */
#define INDIGO_VARIANT_EDGE 1        /* edge-centric, topology-driven: one lane per CSR entry over an edge-source
                                        array built on the device (the suite's E_Topo_Push family) */

/* All resident entries take the device CSR as (nodes, edges, d_nindex,
 * d_nlist[, d_eweight]) and return after the result is in device memory. */
int indigo_bfs(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int src,
               int* d_dist);
int indigo_sssp(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                const int* d_eweight, int src, int* d_dist);
int indigo_cc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label);
/* indigo_cc without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * result as indigo_cc, but the workspace is a fixed array in device memory
 * that the library holds on each device until the process ends (about
 * 64 MiB), so the call makes no hipMalloc / hipFree.  Graphs of at most
 * INDIGO_CC_NOALLOC_MAX_NODES vertices; larger ones get INDIGO_ERR_CAPACITY.
 * All calls on one device share that array: calls that overlap on one
 * device (two contexts, or two host threads) race. */
#define INDIGO_CC_NOALLOC_MAX_NODES 4194304 /* 2^22 */
int indigo_cc_noalloc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist, int* d_label);
int indigo_mis(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist,
               int signed_priority, int* d_status);
/* d_scores: device float[nodes]; initialised to 1/nodes by the call. */
int indigo_pagerank(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist,
                    float* d_scores, int* iterations);
int indigo_tc(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, const int* d_nlist,
              long long* count);

/* ---- batched engine ------------------------------------------------------- */
/* Many graphs of uneven sizes and many sources per graph in ONE persistent
 * launch fed by a device-side task queue (no counterpart in the reference:
 * its harnesses run one query per process).  Results are bit-identical to
 * the resident single-query entries above. */
#define INDIGO_QUERY_BFS 1  /* multi-source BFS: a 64-bit visited mask per vertex */
#define INDIGO_QUERY_SSSP 2 /* multi-source SSSP: a per-source distance slab */
#define INDIGO_QUERY_CC 3
#define INDIGO_QUERY_MIS 4  /* unsigned priority order (the C++ path) */
#define INDIGO_QUERY_PR 5   /* PageRank from 1/nodes, same recurrence and stop rule as indigo_pagerank */
#define INDIGO_MAX_BFS_SOURCES 64
#define INDIGO_MAX_SSSP_SOURCES 32
#define INDIGO_MAX_GANG 256

typedef struct indigo_graph_ref {
  int nodes, edges;
  const int* d_nindex; /* device int[nodes + 1] */
  const int* d_nlist;  /* device int[edges] */
  const int* d_eweight; /* device int[edges]; needed by SSSP queries only */
} indigo_graph_ref;

typedef struct indigo_query {
  int alg;   /* INDIGO_QUERY_* */
  int graph; /* index into the graph table */
  int nsrc;  /* BFS 1..64, SSSP 1..32 sources; ignored for CC / MIS */
  /* workgroups that run this query together: 0 = the engine's rule
   * (indigo_batch_gang), else 1..INDIGO_MAX_GANG (an autotuner's choice);
   * clamped to the launch's grid */
  int gang;
  int src[INDIGO_MAX_BFS_SOURCES];
  /* device output: BFS / SSSP int32[nsrc][nodes], source-major (row i is the
   * distance array of src[i], INT_MAX = unreachable); CC label[nodes]
   * (smallest id of the component); MIS status[nodes] (1 in, 2 out);
   * PR float scores[nodes] followed by the int iteration count (nodes + 1
   * words) */
  int* d_out;
} indigo_query;

/* Runs all queries and returns when every result is in device memory.
 * max_workgroups caps the persistent grid (0: the whole GPU), so that several
 * batches can share the GPU on different contexts. */
int indigo_batch(indigo_ctx* ctx, const indigo_graph_ref* graphs, int ngraphs, const indigo_query* queries,
                 int nqueries, int max_workgroups);

/* The engine's own gang size for a query of `alg` on a graph of this size
 * (before clamping to the grid), or 0 for an unknown algorithm.  Host only. */
int indigo_batch_gang(int alg, int nodes, int edges);

/* ---- graph uploads --------------------------------------------------------- */
/* Copies a host CSR (malloc'd arrays, or an .egr file mapped with
 * indigo_egr_open from indigo_graph.h) into new device arrays through two
 * 64 MiB pinned staging buffers on ctx's stream, with the host-side copy of
 * one chunk overlapping the DMA of the previous one; returns when the graph
 * is resident.  eweight may be NULL.  seconds (may be NULL) receives the
 * wall time of the copies.  Release with indigo_graph_free.  (The reference
 * uploads with blocking cudaMemcpy from malloc'd arrays in each harness,
 * e.g. include/bfs_vertex_cuda.h:143-144.) */
int indigo_graph_upload(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist,
                        const int* eweight, indigo_graph_ref* out, double* seconds);
int indigo_graph_free(indigo_graph_ref* g);

/* d_sp[j] = v for every CSR entry j of vertex v, built on the device (the
 * edge-source array of the edge-form entries; the reference builds it on the
 * host, include/bfs_edge_cuda.h:140-146, tc_edge_cuda.h:159-165). */
int indigo_build_sp(indigo_ctx* ctx, int nodes, int edges, const int* d_nindex, int* d_sp);

/* d_out[v] = d_in[d_perm[v]] for v < n, 4-byte values (no counterpart in the
 * reference): reads a per-vertex result computed on a degree-ordered copy of
 * a graph (indigo_graph_reorder in indigo_graph.h, perm[v] = new id of v)
 * back in the original ids.  d_in and d_out must not overlap.  Asynchronous
 * on the context stream; indigo_ctx_last_ms gives its time. */
int indigo_unpermute(indigo_ctx* ctx, int n, const int* d_perm, const void* d_in, void* d_out);

/* ---- serving ----------------------------------------------------------------- */
/* A request loop over resident graphs (no counterpart in the reference).
 * Requests arrive over time; every free slot (a context with its own stream
 * and a share of the GPU) takes the oldest waiting requests of one cost class
 * (short or long, from the graph's size and the algorithm) that fit the
 * memory budget, merges BFS / SSSP requests on the same graph into
 * multi-source tasks and runs them as one batched launch, while the other
 * slots keep running theirs.  Long work never takes the last free slot, so
 * short requests are not queued behind long launches.  A triangle count runs alone on its slot, and
 * so does an SSSP request on a deep, sparse graph (average degree below 6,
 * at least 65536 vertices), on the single-query near-far kernel. */
#define INDIGO_QUERY_TC 6 /* serving only */

typedef struct indigo_request {
  int alg;   /* INDIGO_QUERY_BFS, _SSSP, _CC, _MIS, _PR or _TC */
  int graph; /* index into the graph table */
  int src;   /* BFS / SSSP source */
  int reserved;
  double arrival; /* seconds after the start of the run */
  /* NULL, or a device buffer for the result: BFS / SSSP / CC / MIS
   * int[nodes], PR float[nodes] then the int iteration count, TC one
   * long long */
  void* d_out;
  /* filled in by indigo_serve, in seconds after the start of the run */
  double admitted; /* taken into a launch */
  double finished; /* result complete (and copied to d_out) */
  int batch;       /* requests served by the same launch */
  int slot;
} indigo_request;

typedef struct indigo_serve_config {
  int device;
  int slots;         /* concurrent launches, 1..16; capped at the process's hardware queues
                        (GPU_MAX_HW_QUEUES, 4 when unset), beyond which streams serialise */
  int max_batch;     /* requests per launch */
  size_t mem_budget; /* device bytes for the outputs of admitted requests */
  double progress_interval; /* > 0: a progress line on stderr every so many seconds */
  /* must be 0.  1 would give the last slot compute units of its own (a
   * CU-masked stream); that mode is switched off and refused with
   * INDIGO_ERR_INVALID_ARG: a run with it ended in a host segmentation fault
   * whose cause is not established */
  int cu_partition;
  /* estimated microseconds of a slot's work one batched launch may hold
   * (0: no cap); a request that would pass it waits for the next launch.  A
   * launch finishes as a whole, so its size bounds the service time of
   * every request in it (caps of 1000 to 4000 us raised p50 and p99 at
   * 150 q/s on the 64-graph pool: the queue grows more than launches
   * shrink) */
  int batch_work_us;
} indigo_serve_config;

typedef struct indigo_serve_stats {
  double makespan;        /* seconds from the start to the last result */
  int launches;           /* batched launches */
  int tc_launches;        /* triangle counts */
  size_t peak_bytes;      /* largest output footprint held at once */
  int max_waiting;        /* longest queue of arrived, unadmitted requests */
  int slots;              /* slots used: cfg.slots capped at the process's hardware queues */
} indigo_serve_stats;

/* Serves reqs (any order of arrival) in real time and returns when every
 * result is complete; per-request times are written back into reqs. */
int indigo_serve(const indigo_graph_ref* graphs, int ngraphs, indigo_request* reqs, int nreqs,
                 const indigo_serve_config* cfg, indigo_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* INDIGO_HIP_H */
