// Triangle counting on gfx950.
//
// Reference semantics: GPUtc_vertex / GPUtc_edge (include/tc_vertex_cuda.h:42,
// include/tc_edge_cuda.h:42) on sorted, self-loop-free lists; h_triCounting
// (tc_vertex_cuda.h:98-117) counts every triangle v < u < w once by sorted
// merges.  The count of a graph does not depend on how its vertices are
// ordered, so any total order counts each triangle exactly once; this file
// orients by (degree, id), which bounds every oriented list by O(sqrt(E)) on
// skewed graphs, and counts in 64 bits (the reference's int overflows on
// large inputs).
//
// Passes (plain launches on the context stream):
//   1. count  : out-degree under the orientation per vertex; lists < kMidDeg
//               by one lane in id order, longer ones binned into mid (one
//               wave each, ballots) and huge (kHugeChunk-entry items, one
//               workgroup each, per-item counts kept); also flags unsorted
//               lists and self loops;
//   2. scan   : exclusive prefix -> oindex;
//   3. fill   : oriented lists, order-preserving (so still sorted), same
//               thread / wave / chunk split as 1;
//   4. count triangles: a wave takes a vertex v (dynamic grabbing), stages
//      N+(v) in LDS (or searches it in global memory when long), and for every
//      u in N+(v) and w in N+(u) - 64 pairs per step with the in-wave
//      balancer - binary-searches w in N+(v).
#include "tc.hpp"

#include "frontier.hpp"
#include "scan.hpp"

namespace indigo {

namespace {

constexpr int kTcBlock = 256;
constexpr int kStage = 1024;  // entries of N+(v) staged in LDS per wave
constexpr int kGrab = 8;      // vertices per dynamic grab

struct TcControl {
  FrontierCounters heavy;  // lists of >= kMidDeg entries
  unsigned long long count;
  unsigned next;
  int error;
  int pad[28];
};

IG_DEV bool outranks(int du, int u, int dv, int v) { return du > dv || (du == dv && u > v); }

IG_DEV int deg(const int* __restrict__ nindex, int v) { return nindex[v + 1] - nindex[v]; }

// Pass 1 / 3 for short lists, one lane per vertex in id order (kFill selects
// the pass); the count pass also bins the long lists.
template <bool kFill>
__global__ __launch_bounds__(kTcBlock) void tc_orient(int nodes, const int* __restrict__ nindex,
                                                      const int* __restrict__ nlist, int* ocount,
                                                      const int* __restrict__ oindex, int* olist, FrontierQueues heavy,
                                                      TcControl* ctl)
{
  const int wpb = kTcBlock / kWave;
  const int nchunks = (nodes + kWave - 1) / kWave;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < nchunks; c += gridDim.x * wpb) {
    const int v = c * kWave + lane_id();
    const bool valid = v < nodes;
    const int beg = valid ? nindex[v] : 0;
    const int end = valid ? nindex[v + 1] : 0;
    const int dv = end - beg;
    if (valid && dv < kMidDeg) {
      int k = 0;
      int bad = 0;
      for (int e = beg; e < end; e++) {
        const int u = nlist[e];
        if (!kFill) bad |= (u == v) | (e > beg && nlist[e - 1] >= u);
        if (outranks(deg(nindex, u), u, dv, v)) {
          if (kFill) olist[oindex[v] + k] = u;
          k++;
        }
      }
      if (!kFill) {
        ocount[v] = k;
        if (bad) st_relaxed(&ctl->error, 1);
      }
    }
    if (!kFill) {
      const bool long_list = valid && dv >= kMidDeg;
      if (long_list) ocount[v] = 0;
      frontier_emit(long_list, v, dv, heavy, &ctl->heavy);
    }
  }
}

// Pass 1 / 3 for long lists: huge ones in kHugeChunk-entry items (one
// workgroup each, per-item counts in hcnt), mid ones one wave each; order is
// kept with ballot prefixes so the oriented lists stay sorted.
template <bool kFill>
__global__ __launch_bounds__(kTcBlock) void tc_orient_heavy(const int* __restrict__ nindex,
                                                            const int* __restrict__ nlist, int* ocount,
                                                            const int* __restrict__ oindex, int* olist,
                                                            FrontierQueues heavy, int* hcnt, TcControl* ctl)
{
  __shared__ int s_owner, s_off;
  __shared__ int s_wcnt[kTcBlock / kWave];
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const unsigned long long hp = ctl->heavy.hpack;
  const unsigned nh = huge_count(hp), nch = huge_chunks(hp);
  for (unsigned item = blockIdx.x; item < nch; item += gridDim.x) {
    if (threadIdx.x == 0) s_owner = huge_owner(heavy.hoff, static_cast<int>(nh), item);
    __syncthreads();
    const int j = s_owner;
    const int v = heavy.hq[j];
    const int vb = nindex[v], ve = nindex[v + 1];
    const int dv = ve - vb;
    const int beg = vb + static_cast<int>(item - heavy.hoff[j]) * kHugeChunk;
    const int end = min(ve, beg + kHugeChunk);
    if (kFill && threadIdx.x == 0) {
      int off = oindex[v];
      for (unsigned k = heavy.hoff[j]; k < item; k++) off += hcnt[k];
      s_off = off;
    }
    __syncthreads();
    int run = kFill ? s_off : 0;
    bool bad = false;
    for (int base = beg; base < end; base += kTcBlock) {
      const int e = base + static_cast<int>(threadIdx.x);
      bool keep = false;
      int u = 0;
      if (e < end) {
        u = nlist[e];
        if (!kFill) bad |= (u == v) || (e > vb && nlist[e - 1] >= u);
        keep = outranks(deg(nindex, u), u, dv, v);
      }
      const unsigned long long m = __ballot(keep);
      if (lane == 0) s_wcnt[wid] = __popcll(m);
      __syncthreads();
      int before = 0, total = 0;
      for (int w = 0; w < kTcBlock / kWave; w++) {
        before += w < wid ? s_wcnt[w] : 0;
        total += s_wcnt[w];
      }
      if (kFill && keep) olist[run + before + lanes_below(m)] = u;
      run += total;
      __syncthreads();
    }
    if (!kFill && threadIdx.x == 0) {
      hcnt[item] = run;
      atomic_add_relaxed(&ocount[v], run);
    }
    if (!kFill && bad) st_relaxed(&ctl->error, 1);
    __syncthreads();
  }
  const unsigned nm = ctl->heavy.nm;
  for (unsigned i = blockIdx.x * (kTcBlock / kWave) + wid; i < nm; i += gridDim.x * (kTcBlock / kWave)) {
    const int v = heavy.mq[i];
    const int vb = nindex[v], ve = nindex[v + 1];
    const int dv = ve - vb;
    int k = kFill ? oindex[v] : 0;
    bool bad = false;
    for (int base = vb; base < ve; base += kWave) {
      const int e = base + lane;
      bool keep = false;
      int u = 0;
      if (e < ve) {
        u = nlist[e];
        if (!kFill) bad |= (u == v) || (e > vb && nlist[e - 1] >= u);
        keep = outranks(deg(nindex, u), u, dv, v);
      }
      const unsigned long long m = __ballot(keep);
      if (kFill && keep) olist[k + lanes_below(m)] = u;
      k += __popcll(m);
    }
    if (!kFill) {
      if (lane == 0) ocount[v] = k;
      if (__ballot(bad) != 0 && lane == 0) st_relaxed(&ctl->error, 1);
    }
  }
}

IG_DEV bool search(const int* list, int n, int key)
{
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (list[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && list[lo] == key;
}

__global__ __launch_bounds__(kTcBlock) void tc_count(int nodes, const int* __restrict__ oindex,
                                                     const int* __restrict__ olist, TcControl* ctl)
{
  __shared__ int s_list[kTcBlock / kWave][kStage];
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  int* stage = s_list[wid];
  unsigned long long cnt = 0;
  for (;;) {
    unsigned start = 0;
    if (lane == 0) start = atomic_add_relaxed(&ctl->next, static_cast<unsigned>(kGrab));
    start = wave_shfl(start, 0);
    if (start >= static_cast<unsigned>(nodes)) break;
    const int stop = min(nodes, static_cast<int>(start) + kGrab);
    for (int v = static_cast<int>(start); v < stop; v++) {
      const int ob = oindex[v], oe = oindex[v + 1];
      const int d = oe - ob;
      if (d < 2) continue;
      const bool staged = d <= kStage;
      if (staged) {
        for (int i = lane; i < d; i += kWave) stage[i] = olist[ob + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      const int* nv = staged ? stage : olist + ob;
      // all pairs (u in N+(v), w in N+(u)), 64 u's per batch, 64 pairs per step
      for (int ub = 0; ub < d; ub += kWave) {
        const int i = ub + lane;
        const int u = i < d ? nv[i] : 0;
        const int b = i < d ? oindex[u] : 0;
        const int len = i < d ? oindex[u + 1] - b : 0;
        const int incl = wave_inclusive_scan(len);
        const int total = wave_shfl(incl, kWave - 1);
        for (int base = 0; base < total; base += kWave) {
          const int k = base + lane;
          const int o = min(wave_owner(incl, k), kWave - 1);
          const int ob2 = wave_shfl(b, o);
          const int oex = wave_shfl(incl - len, o);
          if (k < total) {
            const int w = olist[ob2 + (k - oex)];
            cnt += search(nv, d, w) ? 1ull : 0ull;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt != 0) atomic_add_relaxed(&ctl->count, cnt);
}

int grid_for(long long items, int per_block, int cap)
{
  const long long g = (items + per_block - 1) / per_block;
  return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

}  // namespace

size_t tc_workspace_bytes(int nodes, int edges)
{
  const size_t v = static_cast<size_t>(nodes), e = static_cast<size_t>(edges);
  return Carver::round(sizeof(TcControl)) + 5 * Carver::round((v + 1) * 4) + Carver::round(e * 4 + 64) +
         Carver::round((e / kHugeChunk + e / kHugeDeg + 2) * 4) + Carver::round(scan_temp_bytes(nodes + 1)) + 4096;
}

int tc_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* launches)
{
  Carver cv(ctx->ws);
  TcControl* ctl = cv.take<TcControl>(1);
  int* ocount = cv.take<int>(static_cast<size_t>(nodes) + 1);
  int* oindex = cv.take<int>(static_cast<size_t>(nodes) + 1);
  FrontierQueues heavy{nullptr, cv.take<int>(nodes), cv.take<int>(nodes), cv.take<unsigned>(nodes)};
  int* olist = cv.take<int>(static_cast<size_t>(edges) + 16);
  int* hcnt = cv.take<int>(static_cast<size_t>(edges) / kHugeChunk + static_cast<size_t>(edges) / kHugeDeg + 2);
  int* temp = cv.take<int>(scan_temp_bytes(nodes + 1) / 4 + 1);
  IG_TRY(zero_async(ctx, ctl, sizeof(TcControl)));
  hipStream_t s = ctx->stream;
  const int gw = grid_for((nodes + kWave - 1) / kWave, kTcBlock / kWave, 1 << 16);
  const int gh = ctx->compute_units * 8;
  tc_orient<false><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, ocount, oindex, olist, heavy, ctl);
  tc_orient_heavy<false><<<gh, kTcBlock, 0, s>>>(nindex, nlist, ocount, oindex, olist, heavy, hcnt, ctl);
  IG_TRY(exclusive_scan(s, ocount, oindex, nodes, temp));
  tc_orient<true><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, ocount, oindex, olist, heavy, ctl);
  tc_orient_heavy<true><<<gh, kTcBlock, 0, s>>>(nindex, nlist, ocount, oindex, olist, heavy, hcnt, ctl);
  tc_count<<<ctx->compute_units * 8, kTcBlock, 0, s>>>(nodes, oindex, olist, ctl);
  IG_HIP_TRY(hipGetLastError());
  if (launches) *launches = 10;
  return INDIGO_OK;
}

int tc_collect(indigo_ctx* ctx, long long* count)
{
  Carver cv(ctx->ws);
  TcControl* ctl = cv.take<TcControl>(1);
  IG_TRY(ensure_host_pinned(ctx, sizeof(TcControl)));
  TcControl* h = static_cast<TcControl*>(ctx->host_pinned);
  IG_HIP_TRY(hipMemcpyAsync(h, ctl, sizeof(TcControl), hipMemcpyDeviceToHost, ctx->stream));
  IG_HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (h->error != 0)
    return set_error(INDIGO_ERR_UNSORTED, "ERROR: adjacency list not sorted or contains self edge");
  *count = static_cast<long long>(h->count);
  return INDIGO_OK;
}

}  // namespace indigo
