/*
 * indigo_oracle.c - CPU restatement of the Indigo2 reference algorithms for
 * the six graph kernels.  TEST INFRASTRUCTURE ONLY: it is the checker that
 * the HIP kernels are compared against (tests/, __graft_entry__.smoke() and
 * the CPU-baseline leg of bench.py).  No product code path links or calls it.
 *
 * Every function states the reference file:line it follows.  Parity pins
 * (what each function is checked against, independently of this file):
 *   orc_bfs     golden vectors from the compiled reference (oracle/_ref),
 *               closed form r+c on full grids, brute-force relaxation.
 *   orc_sssp    golden vectors from oracle/_ref, scipy's Dijkstra.
 *   orc_cc      golden vectors from oracle/_ref, scipy's connected
 *               components (labels mapped to component minima).
 *   orc_mis     golden status vectors from oracle/_ref (unsigned order),
 *               independence + maximality, brute-force lexicographic greedy
 *               check on tiny graphs for both orders, reference set sizes.
 *   orc_pr_ref32 iteration counts and scores from oracle/_ref.
 *   orc_pr64    closed forms (regular graph: 1/N; star graph), FP32 ref.
 *   orc_tc      golden counts from oracle/_ref, C(n,3) on cliques, 0 on
 *               grids, brute-force triple enumeration.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF INT_MAX

/* ---------------------------------------------------------------- BFS --
 * Reference: CPUserialDijkstra with unit weights (include/bfs_vertex_cpp.h:52-77,
 * include/bfs_vertex_cuda.h:75-99).  With all weights equal to 1 a FIFO
 * queue settles vertices in non-decreasing distance order, so the distances
 * are those of the priority-queue version. */
void orc_bfs(int nodes, const int* nindex, const int* nlist, int src, int* dist)
{
  int* queue = (int*)malloc(sizeof(int) * (size_t)nodes);
  for (int v = 0; v < nodes; v++) dist[v] = ORC_INF;
  int head = 0, tail = 0;
  dist[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    const int v = queue[head++];
    for (int i = nindex[v]; i < nindex[v + 1]; i++) {
      const int n = nlist[i];
      if (dist[n] == ORC_INF) {
        dist[n] = dist[v] + 1;
        queue[tail++] = n;
      }
    }
  }
  free(queue);
}

/* --------------------------------------------------------------- SSSP --
 * Reference: CPUserialDijkstra (include/sssp_vertex_cpp.h:88-113,
 * include/sssp_vertex_cuda.h:74-99): dist = INT_MAX, dist[src] = 0, relax
 * d = dist[v] + w in int arithmetic.  Binary min-heap of (dist, vertex) with
 * lazy deletion; stale entries are skipped. */
typedef struct {
  long long key; /* dist << 32 | vertex */
} heap_item;

static void heap_push(heap_item* h, long long* n, long long key)
{
  long long i = (*n)++;
  while (i > 0) {
    const long long p = (i - 1) / 2;
    if (h[p].key <= key) break;
    h[i] = h[p];
    i = p;
  }
  h[i].key = key;
}

static long long heap_pop(heap_item* h, long long* n)
{
  const long long top = h[0].key;
  const long long last = h[--(*n)].key;
  long long i = 0;
  for (;;) {
    long long c = 2 * i + 1;
    if (c >= *n) break;
    if (c + 1 < *n && h[c + 1].key < h[c].key) c++;
    if (h[c].key >= last) break;
    h[i] = h[c];
    i = c;
  }
  if (*n > 0) h[i].key = last;
  return top;
}

void orc_sssp(int nodes, const int* nindex, const int* nlist, const int* eweight, int src, int* dist)
{
  const long long cap = (long long)nindex[nodes] + 2;
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (size_t)cap);
  long long hn = 0;
  for (int v = 0; v < nodes; v++) dist[v] = ORC_INF;
  dist[src] = 0;
  heap_push(h, &hn, (long long)src);
  while (hn > 0) {
    const long long key = heap_pop(h, &hn);
    const int v = (int)(key & 0xffffffffLL);
    const int dv = (int)(key >> 32);
    if (dv != dist[v]) continue; /* stale */
    for (int i = nindex[v]; i < nindex[v + 1]; i++) {
      const int n = nlist[i];
      const int d = dv + eweight[i];
      if (d < dist[n]) {
        dist[n] = d;
        heap_push(h, &hn, ((long long)d << 32) | (unsigned)n);
      }
    }
  }
  free(h);
}

/* ----------------------------------------------------------------- CC --
 * Reference: label[v] = v, then min-label propagation to a fixed point
 * (codes/cuda/cc-cuda/CC_CUDA_V_Topo_Push_NonDeterm_IntType_ReadModifyWrite_
 * NonPersist_Atomic_Thread_NoBoundsBug2_NoBoundsBug1.cu:47-76).  On a
 * symmetric graph the fixed point labels every vertex with the smallest id
 * of its component: scanning ids upward, each unlabeled vertex is the
 * minimum of its component and floods it. */
void orc_cc(int nodes, const int* nindex, const int* nlist, int* label)
{
  int* stack = (int*)malloc(sizeof(int) * (size_t)nodes);
  for (int v = 0; v < nodes; v++) label[v] = -1;
  for (int s = 0; s < nodes; s++) {
    if (label[s] >= 0) continue;
    int top = 0;
    label[s] = s;
    stack[top++] = s;
    while (top > 0) {
      const int v = stack[--top];
      for (int i = nindex[v]; i < nindex[v + 1]; i++) {
        const int n = nlist[i];
        if (label[n] < 0) {
          label[n] = s;
          stack[top++] = n;
        }
      }
    }
  }
  free(stack);
}

/* ---------------------------------------------------------------- MIS --
 * Reference: priority[v] = hash(v + 712313887) with the xor-shift-multiply
 * hash of include/mis_vertex_cpp.h:135-140 (= mis_vertex_cuda.h:174-179);
 * an undecided vertex joins the set when every neighbour is excluded or has
 * a smaller (priority, id) pair, and its neighbours are excluded
 * (codes/cpp/mis-cpp/MIS_CPP_V_Topo_Push_IntType_Atomic_CopyDeterm_Block.cpp:
 * 38,51-76 with unsigned priorities; the CUDA variants compare signed ints,
 * MIS_CUDA_V_Topo_Push_Determ_IntType_NonPersist_Atomic_Thread.cu:39-40,66).
 * The fixed point of that rule is the greedy MIS taken in descending
 * (priority, id) order, which this function computes directly.
 * Status codes: 1 included, 2 excluded (mis_vertex_cuda.h:50-52).
 * Self loops are ignored (a vertex is not its own neighbour here). */
unsigned int orc_mis_hash(unsigned int val)
{
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  val = ((val >> 16) ^ val) * 0x45d9f3bu;
  return (val >> 16) ^ val;
}

static int g_mis_signed;
static const unsigned int* g_mis_prio;

static int mis_cmp_desc(const void* pa, const void* pb)
{
  const int a = *(const int*)pa, b = *(const int*)pb;
  const unsigned int ua = g_mis_prio[a], ub = g_mis_prio[b];
  int gt, lt;
  if (g_mis_signed) {
    gt = (int)ua > (int)ub;
    lt = (int)ua < (int)ub;
  } else {
    gt = ua > ub;
    lt = ua < ub;
  }
  if (gt) return -1;
  if (lt) return 1;
  return (a > b) ? -1 : (a < b) ? 1 : 0;
}

void orc_mis(int nodes, const int* nindex, const int* nlist, int signed_priority, int* status)
{
  unsigned int* prio = (unsigned int*)malloc(sizeof(unsigned int) * (size_t)nodes);
  int* order = (int*)malloc(sizeof(int) * (size_t)nodes);
  for (int v = 0; v < nodes; v++) {
    prio[v] = orc_mis_hash((unsigned int)v + 712313887u);
    order[v] = v;
    status[v] = 0;
  }
  g_mis_signed = signed_priority;
  g_mis_prio = prio;
  qsort(order, (size_t)nodes, sizeof(int), mis_cmp_desc);
  for (int k = 0; k < nodes; k++) {
    const int v = order[k];
    if (status[v] != 0) continue;
    status[v] = 1;
    for (int i = nindex[v]; i < nindex[v + 1]; i++)
      if (nlist[i] != v) status[nlist[i]] = 2;
  }
  free(order);
  free(prio);
}

/* ----------------------------------------------------------------- PR --
 * Reference recurrence (include/pr_cuda.h:42-45,126-158;
 * codes/cuda/pr-cuda/PR_CUDA_Pull_FloatType_NonPersist_Thread_Determ_GlobalAdd.cu:
 * 41-69,91-106): s0 = 1/N; each iteration contrib[u] = s[u] / deg[u] and
 * s'[v] = (1 - 0.85) / N + 0.85 * sum_{u in N(v)} contrib[u] (Jacobi, no
 * dangling redistribution); stop when the L1 change of the iteration is
 * <= 1e-4 (the CUDA loop continues while diff > EPSILON) or after 100
 * iterations.
 *
 * orc_pr_ref32 keeps the reference's float arithmetic for scores, contribs
 * and sums (vertex order, sequential), and accumulates the L1 change in
 * double as the reference's C++ variants do
 * (codes/cpp/pr-cpp/PR_CPP_Pull_FloatType_Reduction_Block_Determin.cpp:43-56).
 * Returns the iteration count. */
int orc_pr_ref32(int nodes, const int* nindex, const int* nlist, int max_iter, double eps, float* scores)
{
  float* contrib = (float*)malloc(sizeof(float) * (size_t)nodes);
  const float kDamp = 0.85f;
  const float base = (1.0f - kDamp) / (float)nodes;
  for (int v = 0; v < nodes; v++) scores[v] = 1.0f / (float)nodes;
  int iter = 0;
  double diff;
  do {
    iter++;
    diff = 0.0;
    for (int v = 0; v < nodes; v++) contrib[v] = scores[v] / (float)(nindex[v + 1] - nindex[v]);
    for (int v = 0; v < nodes; v++) {
      float total = 0.0f;
      for (int i = nindex[v]; i < nindex[v + 1]; i++) total += contrib[nlist[i]];
      const float old = scores[v];
      scores[v] = base + kDamp * total;
      diff += fabs((double)scores[v] - (double)old);
    }
  } while (diff > eps && iter < max_iter);
  free(contrib);
  return iter;
}

/* The same recurrence in double precision for exactly `iters` iterations
 * (the FP64 oracle of the GPU's FP32 PageRank).  Returns the L1 change of
 * the last iteration. */
double orc_pr64(int nodes, const int* nindex, const int* nlist, int iters, double* scores)
{
  double* contrib = (double*)malloc(sizeof(double) * (size_t)nodes);
  const double kDamp = 0.85;
  const double base = (1.0 - kDamp) / (double)nodes;
  double diff = 0.0;
  for (int v = 0; v < nodes; v++) scores[v] = 1.0 / (double)nodes;
  for (int it = 0; it < iters; it++) {
    diff = 0.0;
    for (int v = 0; v < nodes; v++) {
      const int d = nindex[v + 1] - nindex[v];
      contrib[v] = d > 0 ? scores[v] / (double)d : 0.0;
    }
    for (int v = 0; v < nodes; v++) {
      double total = 0.0;
      for (int i = nindex[v]; i < nindex[v + 1]; i++) total += contrib[nlist[i]];
      const double s = base + kDamp * total;
      diff += fabs(s - scores[v]);
      scores[v] = s;
    }
  }
  free(contrib);
  return diff;
}

/* ----------------------------------------------------------------- TC --
 * Reference: h_triCounting and h_common (include/tc_vertex_cpp.h:46-65,88-116,
 * = tc_vertex_cuda.h:67-117): for each v and each neighbour u > v at
 * position j, count the common entries of nlist[j+1 .. end(v)) and of u's
 * neighbours greater than u, by a sorted merge.  Each triangle v < u < w is
 * counted once.  Counted here in 64 bits (the reference's basic_t is int). */
static long long orc_common(int beg1, int end1, int beg2, int end2, const int* nlist)
{
  long long c = 0;
  int p1 = beg1, p2 = beg2;
  while (p1 < end1 && p2 < end2) {
    const int a = nlist[p1], b = nlist[p2];
    if (a < b) p1++;
    else if (b < a) p2++;
    else {
      c++;
      p1++;
      p2++;
    }
  }
  return c;
}

unsigned long long orc_tc(int nodes, const int* nindex, const int* nlist)
{
  unsigned long long count = 0;
  for (int v = 0; v < nodes; v++) {
    const int beg1 = nindex[v], end1 = nindex[v + 1];
    int start1 = end1;
    while (beg1 < start1 && v < nlist[start1 - 1]) start1--;
    for (int j = start1; j < end1; j++) {
      const int u = nlist[j];
      const int beg2 = nindex[u], end2 = nindex[u + 1];
      int start2 = end2;
      while (beg2 < start2 && u < nlist[start2 - 1]) start2--;
      count += (unsigned long long)orc_common(j + 1, end1, start2, end2, nlist);
    }
  }
  return count;
}
