// PageRank (FP32 scores, FP64 convergence test) on gfx950.
//
// Reference semantics: PR_GPU (include/pr_cuda.h:48; constants :42-45) with the
// Jacobi pull of codes/cuda/pr-cuda/PR_CUDA_Pull_FloatType_NonPersist_Thread_
// Determ_GlobalAdd.cu:41-69,91-106: contrib[u] = s[u] / deg[u];
// s'[v] = (1 - 0.85) / N + 0.85 * sum_{u in N(v)} contrib[u]; repeat while the
// L1 change of the iteration is > 1e-4 and fewer than 100 iterations ran.
//
// Design: one persistent kernel; per iteration
//   every vertex with a short list (< kPrWaveDeg entries) is summed by one
//   lane in vertex order (coalesced score / contrib / index traffic);
//   longer lists, collected once into a static list, one wave each; lists of
//   kHugeDeg or more in kHugeChunk-entry items, one workgroup each, whose
//   partial sums go to a per-item slot; the workgroup that finishes a huge
//   vertex's last item (a counter per vertex) combines its item partials in
//   item order, so no second barrier is needed;
// then one barrier and the stop test.  Per-vertex sums are FP32 with Neumaier
// compensation (the improved Kahan-Babuska sum: a running error term added
// once at the end); wave and workgroup partials are combined pairwise with the
// same error-free step.  contrib is double-buffered and written by the vertex
// that owns it in the pass that computes its score (Jacobi, as the
// reference's Determ variants), so no separate contrib sweep runs.
// |s' - s| is exact in FP64 and summed in FP64, the only FP64 work: each
// workgroup stores its partials in slots (double-buffered by iteration
// parity); after the barrier every workgroup adds all slots in one fixed
// order, so all compute the bit-identical total and take the same stop
// decision (no floating-point atomics).
#include "pr.hpp"

#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "csum.hpp"
#include "frontier.hpp"
#include "trace.hpp"

namespace indigo {

namespace {

constexpr int kPrBlock = 512;
constexpr int kPrWaveDeg = 16;
constexpr float kDamp = 0.85f;

constexpr int kPrMaxGrid = 4096;
constexpr int kPrUnroll = 4;  // gathers in flight per lane

struct PrControl {
  GridBarrier bar;
  FrontierCounters heavy;
  int iterations;
  int status;
  double last_diff;
  int grid;
  int pad[25];
  // workgroup 0's timeline: entries 0-2 the end of the set-up (scores,
  // heavy lists, first barrier); then per sweep the start and the ends of its
  // huge, mid and short phases (entry 3 + 4 (it - 1) + phase, mode = phase)
  LevelTrace trace;
};

struct PrArgs {
  int nodes;
  int init_uniform;
  int max_iter;
  double eps;
  const int* __restrict__ nindex;
  const int* __restrict__ nlist;
  const int* __restrict__ degree;  // optional caller-provided degrees (PR_GPU contract)
  float* scores;
  float* contrib[2];
  FrontierQueues heavy;  // lists of >= kPrWaveDeg entries (mid) and >= kHugeDeg (huge)
  float2* hpart;         // per huge item: compensated partial (sum, error)
  int hcap;              // huge vertices at most: edges / kHugeDeg + 1
  unsigned* hdone;       // per huge vertex: items finished this sweep
  double* hres;          // per huge vertex: |s' - s| of this sweep
  double* partial;       // [2][2 * kPrMaxGrid] per-workgroup residuals (phase A, phase B)
  int* exit_it;     // [kPrMaxGrid] iteration at which each workgroup left
  double* exit_total;
  PrControl* ctl;
  long long budget;
  long long start_ticks;
};

IG_DEV int deg_of(const PrArgs& a, int v)
{
  return a.degree ? a.degree[v] : a.nindex[v + 1] - a.nindex[v];
}

// acc += cin[nlist[e]] for e = beg + first, beg + first + step, ... < end,
// in that order, with kPrUnroll independent loads in flight per lane: the
// gathers are latency-bound chains (index, then contribution), so a plain
// loop would wait out one round trip per window.
template <int kStep>
IG_DEV void gather(CSum& acc, const float* __restrict__ cin, const int* __restrict__ nlist, int beg, int end,
                   int first)
{
  int e = beg + first;
  for (; e + (kPrUnroll - 1) * kStep < end; e += kPrUnroll * kStep) {
    int u[kPrUnroll];
#pragma unroll
    for (int q = 0; q < kPrUnroll; q++) u[q] = nlist[e + q * kStep];
    float c[kPrUnroll];
#pragma unroll
    for (int q = 0; q < kPrUnroll; q++) c[q] = cin[u[q]];
#pragma unroll
    for (int q = 0; q < kPrUnroll; q++) acc.add(c[q]);
  }
  for (; e < end; e += kStep) acc.add(cin[nlist[e]]);
}

// New score of v from its gathered sum; writes score and next contrib and
// returns |s' - s| in FP64.
IG_DEV double finish_vertex(const PrArgs& a, int v, float total, float base, float* cout)
{
  const float old = a.scores[v];
  const float s = base + kDamp * total;
  a.scores[v] = s;
  cout[v] = s / static_cast<float>(deg_of(a, v));
  return fabs(static_cast<double>(s) - static_cast<double>(old));
}

IG_DEV void st_float2(float2* p, float2 v)
{
  unsigned long long bits;
  __builtin_memcpy(&bits, &v, sizeof(bits));
  st_relaxed(reinterpret_cast<unsigned long long*>(p), bits);
}

IG_DEV float2 ld_float2(const float2* p)
{
  const unsigned long long bits = ld_relaxed(reinterpret_cast<const unsigned long long*>(p));
  float2 v;
  __builtin_memcpy(&v, &bits, sizeof(v));
  return v;
}

// Workgroup-wide compensated sum of one partial per thread; result valid in
// thread 0.
IG_DEV CSum block_merge(CSum acc, float* s_ps, float* s_pc)
{
  acc = wave_merge(acc);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) {
    s_ps[wid] = acc.s;
    s_pc[wid] = acc.c;
  }
  __syncthreads();
  CSum blk;
  if (threadIdx.x == 0)
    for (int w = 0; w < kPrBlock / kWave; w++) blk.merge(s_ps[w], s_pc[w]);
  __syncthreads();
  return blk;
}

// FP64 sum of one value per thread; result valid in every thread.
IG_DEV double block_sum_f64(double x, double* s_red)
{
  x = wave_sum(x);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) s_red[wid] = x;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < kPrBlock / kWave; w++) t += s_red[w];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(kPrBlock) void pr_kernel(PrArgs a)
{
  __shared__ double s_red[kPrBlock / kWave];
  __shared__ float s_ps[kPrBlock / kWave], s_pc[kPrBlock / kWave];
  __shared__ int s_owner;
  __shared__ int s_last;
  __shared__ HeavyCollector s_emit;
  PrControl* ctl = a.ctl;
  if (!residency_barrier(&ctl->bar, a.start_ticks)) return;
  const long long deadline = now_ticks() + a.budget;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = gridDim.x * blockDim.x;
  const int lane = lane_id();
  const int wpb = kPrBlock / kWave;
  const float inv_n = 1.0f / static_cast<float>(a.nodes);
  const float base = (1.0f - kDamp) / static_cast<float>(a.nodes);

  if (tid == 0) ctl->trace.start = now_ticks();
  for (int v = tid; v < a.nodes; v += stride) {
    const float s = a.init_uniform ? inv_n : a.scores[v];
    a.scores[v] = s;
    a.contrib[0][v] = s / static_cast<float>(deg_of(a, v));
  }
  for (int j = tid; j < a.hcap; j += stride) st_relaxed(&a.hdone[j], 0u);
  trace_level(&ctl->trace, 0, 0u, 0ull, 0);
  collect_heavy<kPrWaveDeg>(s_emit, a.nodes, a.nindex, a.heavy, &ctl->heavy);
  trace_level(&ctl->trace, 1, 0u, 0ull, 1);
  if (!grid_barrier(&ctl->bar, deadline)) {
    if (tid == 0) ctl->status = 1;
    if (threadIdx.x == 0) a.exit_it[blockIdx.x] = -1000000;
    return;
  }
  trace_level(&ctl->trace, 2, 0u, 0ull, 2);
  const unsigned nm = ld_relaxed(&ctl->heavy.nm);
  const unsigned long long hp = ld_relaxed(&ctl->heavy.hpack);
  const unsigned nh = huge_count(hp), nch = huge_chunks(hp);

  int it = 0;
  for (;;) {
    it++;
    const float* cin = a.contrib[(it - 1) & 1];
    float* cout = a.contrib[it & 1];
    double* slots = a.partial + (it & 1) * (2 * kPrMaxGrid);
    double d = 0.0;
    const int tl = 3 + 4 * (it - 1);
    trace_level(&ctl->trace, tl, static_cast<unsigned>(it), 0ull, 0);

    // huge items -> partials; the last item of a vertex combines them
    for (unsigned item = blockIdx.x; item < nch; item += gridDim.x) {
      if (threadIdx.x == 0) s_owner = huge_owner(a.heavy.hoff, static_cast<int>(nh), item);
      __syncthreads();
      const int j = s_owner;
      const int v = a.heavy.hq[j];
      const int len = a.nindex[v + 1] - a.nindex[v];
      const int items = (len + kHugeChunk - 1) / kHugeChunk;
      const unsigned first = a.heavy.hoff[j];
      const int beg = a.nindex[v] + static_cast<int>(item - first) * kHugeChunk;
      const int end = min(a.nindex[v + 1], beg + kHugeChunk);
      CSum acc;
      gather<kPrBlock>(acc, cin, a.nlist, beg, end, static_cast<int>(threadIdx.x));
      const CSum blk = block_merge(acc, s_ps, s_pc);
      if (threadIdx.x == 0) {
        // the partial and the counter are device-scope accesses, the store
        // completed before the count (agent-scope acquire / release fences
        // would write back and invalidate the XCD's L2, and with it the
        // cached contributions every gather reads: 6.8 -> 9.7 ms)
        st_float2(&a.hpart[item], make_float2(blk.s, blk.c));
        wave_drain();
        const unsigned k = atomic_add_relaxed(&a.hdone[j], 1u);
        s_last = k + 1 == static_cast<unsigned>(items);
      }
      __syncthreads();
      if (s_last && threadIdx.x < kWave) {
        CSum tot;
        for (int k = lane; k < items; k += kWave) {
          const float2 p = ld_float2(&a.hpart[first + k]);
          tot.merge(p.x, p.y);
        }
        tot = wave_merge(tot);
        if (lane == 0) {
          st_relaxed(&a.hres[j], finish_vertex(a, v, tot.value(), base, cout));
          st_relaxed(&a.hdone[j], 0u);  // every item of this sweep has counted
        }
      }
      __syncthreads();
    }
    trace_level(&ctl->trace, tl + 1, static_cast<unsigned>(it), 0ull, 1);
    // phase A: mid lists, one wave each
    for (unsigned i = blockIdx.x * wpb + threadIdx.x / kWave; i < nm; i += gridDim.x * wpb) {
      const int v = a.heavy.mq[i];
      const int beg = a.nindex[v], end = a.nindex[v + 1];
      CSum acc;
      gather<kWave>(acc, cin, a.nlist, beg, end, lane);
      acc = wave_merge(acc);
      if (lane == 0) d += finish_vertex(a, v, acc.value(), base, cout);
    }
    trace_level(&ctl->trace, tl + 2, static_cast<unsigned>(it), 0ull, 2);
    // phase A: short lists, one lane each, vertex order
    for (int v = tid; v < a.nodes; v += stride) {
      const int beg = a.nindex[v], end = a.nindex[v + 1];
      if (end - beg >= kPrWaveDeg) continue;
      CSum acc;
      gather<1>(acc, cin, a.nlist, beg, end, 0);
      d += finish_vertex(a, v, acc.value(), base, cout);
    }
    d = block_sum_f64(d, s_red);
    if (threadIdx.x == 0) slots[2 * blockIdx.x] = d;
    d = 0.0;
    trace_level(&ctl->trace, tl + 3, static_cast<unsigned>(it), 0ull, 3);

    if (!grid_barrier(&ctl->bar, deadline)) {
      if (tid == 0) ctl->status = 1;
      if (threadIdx.x == 0) a.exit_it[blockIdx.x] = -it;
      return;
    }
    // every workgroup adds all slots, then the huge vertices' residuals, in
    // one fixed order: the same total everywhere, whichever workgroup
    // combined which huge vertex
    double part = 0.0;
    for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += kPrBlock) part += slots[2 * b];
    double hpart_sum = 0.0;
    for (int j = threadIdx.x; j < static_cast<int>(nh); j += kPrBlock) hpart_sum += ld_relaxed(&a.hres[j]);
    const double total = block_sum_f64(part, s_red) + block_sum_f64(hpart_sum, s_red);
    if (!(total > a.eps && it < a.max_iter)) {
      if (tid == 0) {
        ctl->iterations = it;
        ctl->last_diff = total;
        ctl->grid = gridDim.x;
      }
      if (threadIdx.x == 0) {
        a.exit_it[blockIdx.x] = it;
        a.exit_total[blockIdx.x] = total;
      }
      break;
    }
  }
}

}  // namespace

// Upper bound on the number of huge items: every huge list has at least
// kHugeDeg entries, so items <= edges / kHugeChunk + edges / kHugeDeg.
static size_t pr_max_chunks(int edges)
{
  return static_cast<size_t>(edges) / kHugeChunk + static_cast<size_t>(edges) / kHugeDeg + 2;
}

size_t pr_workspace_bytes(int nodes, int edges)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(PrControl)) + 6 * Carver::round(v * 4) + Carver::round(pr_max_chunks(edges) * 8) +
         Carver::round((static_cast<size_t>(edges) / kHugeDeg + 1) * 12) + 256 +
         Carver::round(4 * kPrMaxGrid * 8) + Carver::round(kPrMaxGrid * 4) + Carver::round(kPrMaxGrid * 8) + 4096;
}

// Grid of one call: a sweep's barrier and combine cost more than a small
// graph's gathers, so the grid grows with the graph, as the power of two
// nearest 0.2 * sqrt(E) workgroups, at least 64.  Measured on R-MAT 12, 14, 16
// and 18 (0.1M to 7.6M entries) the best caps were 64, 128, 256 and 512
// (calls 0.31 -> 0.15, 0.35 -> 0.19, 0.45 -> 0.31 and 0.69 -> 0.64 ms); from
// about 2^25 entries on the rule exceeds the resident grid, which is kept.
static int pr_grid_cap(int edges)
{
  const double want = 0.2 * std::sqrt(static_cast<double>(edges));
  int cap = 64;
  while (cap < kPrMaxGrid && cap * 1.41421356 < want) cap *= 2;
  return cap;
}

int pr_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, const int* degree,
           float* scores, int init_uniform)
{
  Carver cv(ctx->ws);
  PrArgs a{};
  a.ctl = cv.take<PrControl>(1);
  a.contrib[0] = cv.take<float>(nodes);
  a.contrib[1] = cv.take<float>(nodes);
  a.heavy.q = nullptr;
  a.heavy.mq = cv.take<int>(nodes);
  a.heavy.hq = cv.take<int>(nodes);
  a.heavy.hoff = cv.take<unsigned>(nodes);
  a.hpart = cv.take<float2>(pr_max_chunks(edges));
  a.hcap = edges / kHugeDeg + 1;  // every huge list holds kHugeDeg entries or more
  a.hdone = cv.take<unsigned>(static_cast<size_t>(a.hcap));
  a.hres = cv.take<double>(static_cast<size_t>(a.hcap));
  a.partial = cv.take<double>(4 * kPrMaxGrid);
  a.exit_it = cv.take<int>(kPrMaxGrid);
  a.exit_total = cv.take<double>(kPrMaxGrid);
  ctx->diag[0] = a.exit_it;
  ctx->diag[1] = a.exit_total;
  a.nodes = nodes;
  a.init_uniform = init_uniform;
  a.max_iter = 100;
  a.eps = static_cast<double>(0.0001f);
  a.nindex = nindex;
  a.nlist = nlist;
  a.degree = degree;
  a.scores = scores;
  a.budget = budget_ticks(ctx);
  a.start_ticks = residency_ticks(ctx);
  IG_TRY(zero_async(ctx, a.ctl, sizeof(PrControl)));
  IG_TRY(zero_async(ctx, a.exit_it, sizeof(int) * kPrMaxGrid));
  void* args[] = {&a};
  IG_TRY(coop_launch(ctx, reinterpret_cast<const void*>(pr_kernel), kPrBlock, 0, args, pr_grid_cap(edges), nullptr,
                     "pr_kernel"));
  return INDIGO_OK;
}

int pr_collect(indigo_ctx* ctx, int* iterations, double* last_diff)
{
  Carver cv(ctx->ws);
  PrControl* ctl = cv.take<PrControl>(1);
  IG_TRY(ensure_host_pinned(ctx, sizeof(PrControl)));
  PrControl* h = static_cast<PrControl*>(ctx->host_pinned);
  IG_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(PrControl), hipMemcpyDeviceToHost, ctx->stream));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  IG_TRY(check_abort(ctx, reinterpret_cast<const void*>(pr_kernel), h->bar.abort.v, "pagerank"));
  if (h->status != 0) {
    // which iteration each workgroup left at (negative: its barrier gave
    // up), and the stop total it computed
    const int g = kPrMaxGrid;
    std::vector<int> it(g);
    std::vector<double> tot(g);
    IG_HIP_TRY(hipMemcpy(it.data(), ctx->diag[0], sizeof(int) * g, hipMemcpyDeviceToHost));
    IG_HIP_TRY(hipMemcpy(tot.data(), ctx->diag[1], sizeof(double) * g, hipMemcpyDeviceToHost));
    std::map<int, int> by_it;
    std::map<double, int> by_total;
    for (int b = 0; b < g; b++) {
      if (it[b] == 0) continue;
      by_it[it[b]]++;
      if (it[b] > 0) by_total[tot[b]]++;
    }
    std::string msg;
    for (auto& kv : by_it) msg += " it" + std::to_string(kv.first) + "x" + std::to_string(kv.second);
    msg += " | totals:";
    int shown = 0;
    for (auto& kv : by_total) {
      if (shown++ == 6) break;
      char buf[64];
      snprintf(buf, sizeof(buf), " %.17g x%d", kv.first, kv.second);
      msg += buf;
    }
    char bar[160];
    snprintf(bar, sizeof(bar), " | barrier groups %u %u %u %u %u %u %u %u top %u gen %u",
             h->bar.group[0].v, h->bar.group[1].v, h->bar.group[2].v, h->bar.group[3].v, h->bar.group[4].v,
             h->bar.group[5].v, h->bar.group[6].v, h->bar.group[7].v, h->bar.top.v, h->bar.gen.v);
    msg += bar;
    msg += " | grid " + std::to_string(ctx->last_grid) + " per_cu " + std::to_string(ctx->last_per_cu) + " cus " +
           std::to_string(ctx->compute_units);
    return set_error(INDIGO_ERR_TIMEOUT, "pagerank: persistent kernel stopped (status %d); workgroups:%s",
                     h->status, msg.c_str());
  }
  store_trace(ctx, h->trace);
  if (iterations) *iterations = h->iterations;
  if (last_diff) *last_diff = h->last_diff;
  return INDIGO_OK;
}

}  // namespace indigo
