// Triangle counting on gfx950.
//
// Reference semantics: GPUtc_vertex / GPUtc_edge (include/tc_vertex_cuda.h:42,
// include/tc_edge_cuda.h:42) on sorted, self-loop-free lists; h_triCounting
// (tc_vertex_cuda.h:98-117) counts every triangle v < u < w once by sorted
// merges.  The count of a graph does not depend on how its vertices are
// ordered, so any total order counts each triangle exactly once; this file
// orients by (degree, id), which bounds every oriented list N+(v) by
// sqrt(2m) for m undirected edges, and counts in 64 bits (the reference's
// int overflows on large inputs).
//
// Each triangle is the pair (v -> u, w) with u, w in N+(v) and w in N+(u):
// it is found from the oriented edge v -> u by intersecting N+(v) and N+(u).
// The intersection iterates the SHORTER of the two lists and probes the
// other through an LDS hash table; the task of vertex x hashes N+(x) once and
// handles
//   * its out-edges x -> u with |N+(u)| <= |N+(x)| (iterating N+(u)), and
//   * its in-list N-(x) = { v : v -> x, |N+(v)| < |N+(x)| } (iterating N+(v)),
// so every oriented edge is taken exactly once, by its endpoint with the
// longer list (2.4x fewer probes than always iterating N+(u) on R-MAT).
//
// Passes (plain launches on the context stream):
//   1. out-count : |N+(v)| per vertex; lists < kMidDeg by one lane in id
//                  order, longer ones binned into mid (one wave each,
//                  ballots) and huge (kHugeChunk-entry items, one workgroup
//                  each, per-item counts kept); also flags unsorted lists
//                  and self loops;
//   2. scan      : exclusive prefix -> oindex;
//   3. out-fill  : N+ lists, order-preserving (so still sorted), same split;
//   4-6.           the same three steps for the in-lists N-(x);
//   7. count, wave pass: a wave grabs 64 consecutive vertices at a time from
//      a global counter (a static stride would be badly unbalanced: on R-MAT
//      graphs a vertex's degree follows the bit pattern of its id); a task
//      with 2 <= |N+(x)| <= kWaveList and at most kWaveWork probes
//      ((|N+(x)| + |N-(x)|) * |N+(x)| bounds them) builds an open-addressing
//      table of N+(x) in LDS (load factor <= 1/4, Fibonacci hashing, linear
//      probing) and walks the (list, element) pairs 64 per step with the
//      in-wave balancer, one probe each.  Every other task is cut into parts
//      of kPartItems items (out-list, then in-list), queued for pass 8;
//   8. count, workgroup passes: a workgroup grabs a part, hashes N+(x) and
//      spreads the part's items over its waves (512 threads and a 32 KiB
//      table up to kMidList entries, 1024 threads and a 128 KiB table
//      beyond).  Parts bound the work of one grab, so a few heavy vertices
//      cannot form a tail.  A list too long for the table (m > 2^25 edges, or
//      a hint) is searched by binary search in global memory instead.
#include "tc.hpp"

#include "frontier.hpp"
#include "scan.hpp"

namespace indigo {

namespace {

constexpr int kTcBlock = 256;
constexpr int kWaveList = 256;    // longest N+(x) the wave pass takes
constexpr int kWaveSlots = 1024;  // its LDS table (4 slots per entry)
constexpr int kMidBlock = 512;
constexpr int kMidList = 2048;
constexpr int kMidSlots = 8192;
constexpr int kBigBlock = 1024;
constexpr int kBigSlots = 32768;
static_assert(kMidBlock == 512 && kMidSlots == 8192, "the main-kernel name below spells these out");
constexpr int kWaveWork = 1 << 15;  // probe bound of a wave-pass task
constexpr int kPartItems = 1024;    // items per workgroup-pass part
constexpr int kUnroll = 4;          // windows of 64 per step over a long list
constexpr int kLongList = 32;       // lists walked by the whole wave

struct TcControl {
  FrontierCounters heavy;  // lists of >= kMidDeg entries
  unsigned long long count;
  unsigned nmid, nbig;     // tasks deferred to the workgroup passes
  unsigned next, next_mid, next_big;
  int error;
  int pad[24];
};

IG_DEV bool outranks(int du, int u, int dv, int v) { return du > dv || (du == dv && u > v); }

IG_DEV int deg(const int* __restrict__ nindex, int v) { return nindex[v + 1] - nindex[v]; }

// Out-list of v: the neighbours that outrank v.
struct OutSel {
  IG_DEV bool operator()(const int* __restrict__ nindex, int v, int dv, int u) const
  {
    return outranks(deg(nindex, u), u, dv, v);
  }
};

// In-list of v: the neighbours u that v outranks (u -> v) whose out-list is
// shorter than v's.
struct InSel {
  const int* oindex;
  IG_DEV bool operator()(const int* __restrict__ nindex, int v, int dv, int u) const
  {
    if (outranks(deg(nindex, u), u, dv, v)) return false;
    return oindex[u + 1] - oindex[u] < oindex[v + 1] - oindex[v];
  }
};

// Selection over short lists (< kMidDeg), one lane per vertex in id order.
// Count pass (!kFill): count[v] = entries kept (0 for long lists, which the
// heavy pass adds); with kBin it also bins the long lists into `heavy` and
// flags unsorted lists and self loops.  Fill pass: list[index[v] + k].
template <bool kFill, bool kBin, typename Sel>
__global__ __launch_bounds__(kTcBlock) void tc_select(int nodes, const int* __restrict__ nindex,
                                                      const int* __restrict__ nlist, Sel sel, int* count,
                                                      const int* __restrict__ index, int* list, FrontierQueues heavy,
                                                      TcControl* ctl)
{
  __shared__ FrontierEmitter s_emit;
  const int wpb = kTcBlock / kWave;
  const int nchunks = (nodes + kWave - 1) / kWave;
  if (kBin) fe_reset(s_emit);
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < nchunks; c += gridDim.x * wpb) {
    const int v = c * kWave + lane_id();
    const bool valid = v < nodes;
    const int beg = valid ? nindex[v] : 0;
    const int end = valid ? nindex[v + 1] : 0;
    const int dv = end - beg;
    if (valid && dv < kMidDeg) {
      int k = 0;
      int bad = 0;
      const int base = kFill ? index[v] : 0;
      for (int e = beg; e < end; e++) {
        const int u = nlist[e];
        if (kBin) bad |= (u == v) | (e > beg && nlist[e - 1] >= u);
        if (sel(nindex, v, dv, u)) {
          if (kFill) list[base + k] = u;
          k++;
        }
      }
      if (!kFill) count[v] = k;
      if (kBin && bad) st_relaxed(&ctl->error, 1);
    }
    if (!kFill) {
      const bool long_list = valid && dv >= kMidDeg;
      if (long_list) count[v] = 0;
      if (kBin) frontier_emit_block(s_emit, long_list, v, dv, heavy, &ctl->heavy);
    }
  }
  if (kBin) fe_flush(s_emit, heavy, &ctl->heavy);
}

// Selection over the binned long lists: huge ones in kHugeChunk-entry items
// (one workgroup each, per-item counts in hcnt), mid ones one wave each;
// order is kept with ballot prefixes so selected lists stay sorted.
template <bool kFill, bool kCheck, typename Sel>
__global__ __launch_bounds__(kTcBlock) void tc_select_heavy(const int* __restrict__ nindex,
                                                            const int* __restrict__ nlist, Sel sel, int* count,
                                                            const int* __restrict__ index, int* list,
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
      int off = index[v];
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
        if (kCheck) bad |= (u == v) || (e > vb && nlist[e - 1] >= u);
        keep = sel(nindex, v, dv, u);
      }
      const unsigned long long m = __ballot(keep);
      if (lane == 0) s_wcnt[wid] = __popcll(m);
      __syncthreads();
      int before = 0, total = 0;
      for (int w = 0; w < kTcBlock / kWave; w++) {
        before += w < wid ? s_wcnt[w] : 0;
        total += s_wcnt[w];
      }
      if (kFill && keep) list[run + before + lanes_below(m)] = u;
      run += total;
      __syncthreads();
    }
    if (!kFill && threadIdx.x == 0) {
      hcnt[item] = run;
      atomic_add_relaxed(&count[v], run);
    }
    if (kCheck && bad) st_relaxed(&ctl->error, 1);
    __syncthreads();
  }
  const unsigned nm = ctl->heavy.nm;
  for (unsigned i = blockIdx.x * (kTcBlock / kWave) + wid; i < nm; i += gridDim.x * (kTcBlock / kWave)) {
    const int v = heavy.mq[i];
    const int vb = nindex[v], ve = nindex[v + 1];
    const int dv = ve - vb;
    int k = kFill ? index[v] : 0;
    bool bad = false;
    for (int base = vb; base < ve; base += kWave) {
      const int e = base + lane;
      bool keep = false;
      int u = 0;
      if (e < ve) {
        u = nlist[e];
        if (kCheck) bad |= (u == v) || (e > vb && nlist[e - 1] >= u);
        keep = sel(nindex, v, dv, u);
      }
      const unsigned long long m = __ballot(keep);
      if (kFill && keep) list[k + lanes_below(m)] = u;
      k += __popcll(m);
    }
    if (!kFill && lane == 0) count[v] = k;
    if (kCheck && __ballot(bad) != 0 && lane == 0) st_relaxed(&ctl->error, 1);
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

// Hash table geometry for a list of d entries: a power of two >= 4d slots
// (at least 64); the slot of w is the top log2(size) bits of w * 2654435761
// (Fibonacci hashing).  Empty slots hold -1.
struct TcTable {
  int size;
  unsigned mask;
  int shift;
};

IG_DEV TcTable table_for(int d)
{
  const int size = max(64, 1 << (32 - __clz(4 * d - 1)));
  return TcTable{size, static_cast<unsigned>(size - 1), __clz(size) + 1};
}

IG_DEV unsigned slot_of(int w, int shift) { return (static_cast<unsigned>(w) * 2654435761u) >> shift; }

IG_DEV void table_insert(int* tab, const TcTable& t, int w)
{
  unsigned h = slot_of(w, t.shift);
  while (atomicCAS(&tab[h], -1, w) != -1) h = (h + 1) & t.mask;
}

IG_DEV bool table_find(const int* tab, const TcTable& t, int w)
{
  unsigned h = slot_of(w, t.shift);
  int x;
  while ((x = tab[h]) != -1 && x != w) h = (h + 1) & t.mask;
  return x == w;
}

IG_DEV void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The task of vertex x: N+(x) at olist[ob, ob + d), its in-list at
// ilist[ib, ib + di).
struct TcTask {
  int ob, d, ib, di;
};

// Wave-collective: over items[ub0 + 64k + lane] (k = 0, 1, ... stepping by
// ustep), counts the w in N+(item) found in N+(x) - through `tab`, or by
// binary search of the sorted nx[0, d) when tab is null - for the items
// whose list has at most `limit` entries.  Lists of kLongList or more
// entries (97% of the probes on R-MAT) are walked by the whole wave, one at a
// time, kUnroll windows of 64 per step so that their loads overlap; shorter
// ones share windows of 64 pairs through the in-wave balancer.
IG_DEV unsigned long long count_pairs(const int* __restrict__ oindex, const int* __restrict__ olist,
                                      const int* __restrict__ items, int n, int limit, int ub0, int ustep,
                                      const int* tab, const TcTable& t, const int* nx, int d)
{
  const int lane = lane_id();
  unsigned long long cnt = 0;
  auto probe = [&](int w) { return tab != nullptr ? table_find(tab, t, w) : search(nx, d, w); };
  for (int ub = ub0; ub < n; ub += ustep) {
    const int i = ub + lane;
    const int u = i < n ? items[i] : 0;
    const int b = i < n ? oindex[u] : 0;
    const int l = i < n ? oindex[u + 1] - b : 0;
    const int len = l <= limit ? l : 0;
    unsigned long long todo = __ballot(len >= kLongList);
    while (todo != 0) {
      const int o = __ffsll(static_cast<long long>(todo)) - 1;
      todo &= todo - 1;
      const int ob2 = wave_bcast(b, o), ol = wave_bcast(len, o);
      for (int base = 0; base < ol; base += kUnroll * kWave) {
        int w[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; q++) {
          const int k = base + q * kWave + lane;
          w[q] = k < ol ? olist[ob2 + k] : -1;
        }
#pragma unroll
        for (int q = 0; q < kUnroll; q++)
          if (w[q] >= 0) cnt += probe(w[q]) ? 1ull : 0ull;
      }
    }
    const int slen = len < kLongList ? len : 0;
    const int incl = wave_inclusive_scan(slen);
    const int total = wave_bcast(incl, kWave - 1);
    const int shift = b - (incl - slen);  // pair k is olist[shift[o] + k] for its owner lane o
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int at = wave_shfl(shift, o) + k;
      if (k < total) cnt += probe(olist[at]) ? 1ull : 0ull;
    }
  }
  return cnt;
}

// Items [lo, hi) of the task's item sequence (N+(x), then N-(x)).
IG_DEV unsigned long long task_pairs(const int* __restrict__ oindex, const int* __restrict__ olist,
                                     const int* __restrict__ ilist, const TcTask& x, int lo, int hi, int ub0,
                                     int ustep, const int* tab, const TcTable& t)
{
  const int* nx = olist + x.ob;
  unsigned long long cnt = 0;
  const int a1 = min(hi, x.d);
  if (lo < a1) cnt += count_pairs(oindex, olist, nx + lo, a1 - lo, x.d, ub0, ustep, tab, t, nx, x.d);
  const int b0 = max(lo, x.d) - x.d, b1 = hi - x.d;
  if (b0 < b1) cnt += count_pairs(oindex, olist, ilist + x.ib + b0, b1 - b0, x.d, ub0, ustep, tab, t, nx, x.d);
  return cnt;
}

// Wave-collective: appends the parts (x << 32 | part) of each lane's task.
IG_DEV void append_parts(int parts, int x, unsigned long long* queue, unsigned* count)
{
  const int incl = wave_inclusive_scan(parts);
  const int total = wave_shfl(incl, kWave - 1);
  if (total == 0) return;
  unsigned base = 0;
  if (lane_id() == 0) base = atomic_add_relaxed(count, static_cast<unsigned>(total));
  base = wave_shfl(base, 0);
  for (int q = 0; q < parts; q++)
    queue[base + static_cast<unsigned>(incl - parts + q)] =
        (static_cast<unsigned long long>(x) << 32) | static_cast<unsigned>(q);
}

__global__ __launch_bounds__(kTcBlock) void tc_count(int nodes, const int* __restrict__ oindex,
                                                     const int* __restrict__ olist, const int* __restrict__ iindex,
                                                     const int* __restrict__ ilist, unsigned long long* mid,
                                                     unsigned long long* big, TcControl* ctl)
{
  __shared__ int s_tab[kTcBlock / kWave][kWaveSlots];
  const int lane = lane_id();
  int* tab = s_tab[threadIdx.x / kWave];
  unsigned long long cnt = 0;
  for (;;) {
    unsigned v0 = 0;
    if (lane == 0) v0 = atomic_add_relaxed(&ctl->next, static_cast<unsigned>(kWave));
    v0 = wave_shfl(v0, 0);
    if (v0 >= static_cast<unsigned>(nodes)) break;
    // 64 consecutive vertices per grab; heavy tasks go to the workgroup
    // passes in parts (|N+(x)| < 2: no triangle can be found through N+(x))
    const int vl = static_cast<int>(v0) + lane;
    const int dl = vl < nodes ? oindex[vl + 1] - oindex[vl] : 0;
    const int il = vl < nodes ? iindex[vl + 1] - iindex[vl] : 0;
    const bool here = dl >= 2 && dl <= kWaveList && (dl + il) * dl <= kWaveWork;
    const int parts = dl >= 2 && !here ? (dl + il + kPartItems - 1) / kPartItems : 0;
    append_parts(dl <= kMidList ? parts : 0, vl, mid, &ctl->nmid);
    append_parts(dl > kMidList ? parts : 0, vl, big, &ctl->nbig);
    unsigned long long todo = __ballot(here);
    while (todo != 0) {
      const int j = __ffsll(static_cast<long long>(todo)) - 1;
      todo &= todo - 1;
      const int v = static_cast<int>(v0) + j;
      const TcTask x{oindex[v], wave_shfl(dl, j), iindex[v], wave_shfl(il, j)};
      const TcTable t = table_for(x.d);
      for (int i = lane; i < t.size; i += kWave) tab[i] = -1;
      wave_lds_sync();
      for (int i = lane; i < x.d; i += kWave) table_insert(tab, t, olist[x.ob + i]);
      wave_lds_sync();
      cnt += task_pairs(oindex, olist, ilist, x, 0, x.d + x.di, 0, kWave, tab, t);
      wave_lds_sync();
    }
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt != 0) atomic_add_relaxed(&ctl->count, cnt);
}

// One workgroup per part from `parts` (grabbed through *next); lists up to
// lds_list entries (<= kSlots / 4) are hashed into LDS.
template <int kBlock, int kSlots>
__global__ __launch_bounds__(kBlock) void tc_count_group(const int* __restrict__ oindex,
                                                         const int* __restrict__ olist,
                                                         const int* __restrict__ iindex,
                                                         const int* __restrict__ ilist,
                                                         const unsigned long long* __restrict__ parts,
                                                         const unsigned* nparts, unsigned* next, int lds_list,
                                                         TcControl* ctl)
{
  __shared__ int s_tab[kSlots];
  __shared__ unsigned s_k;
  constexpr int kWaves = kBlock / kWave;
  const int wid = threadIdx.x / kWave;
  const unsigned n = *nparts;
  unsigned long long cnt = 0;
  for (;;) {
    if (threadIdx.x == 0) s_k = atomic_add_relaxed(next, 1u);
    __syncthreads();
    const unsigned k = s_k;
    if (k >= n) break;
    const unsigned long long p = parts[k];
    const int v = static_cast<int>(p >> 32);
    const int lo = static_cast<int>(p & 0xffffffffull) * kPartItems;
    const int ob = oindex[v], ib = iindex[v];
    const TcTask x{ob, oindex[v + 1] - ob, ib, iindex[v + 1] - ib};
    const bool hashed = x.d <= lds_list;
    const TcTable t = table_for(hashed ? x.d : 1);
    if (hashed) {
      for (int i = threadIdx.x; i < t.size; i += kBlock) s_tab[i] = -1;
      __syncthreads();
      for (int i = threadIdx.x; i < x.d; i += kBlock) table_insert(s_tab, t, olist[x.ob + i]);
      __syncthreads();
    }
    const int hi = min(lo + kPartItems, x.d + x.di);
    cnt += task_pairs(oindex, olist, ilist, x, lo, hi, wid * kWave, kWaves * kWave, hashed ? s_tab : nullptr, t);
    __syncthreads();
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt != 0) atomic_add_relaxed(&ctl->count, cnt);
}

int grid_for(long long items, int per_block, int cap)
{
  const long long g = (items + per_block - 1) / per_block;
  return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

size_t list_words(int edges) { return static_cast<size_t>(edges) + 16; }

// Parts: at most one per vertex plus one per kPartItems list entries.
size_t part_words(int nodes, int edges)
{
  return static_cast<size_t>(nodes) + 2 * (static_cast<size_t>(edges) / kPartItems) + 64;
}

size_t hcnt_words(int edges)
{
  const size_t e = static_cast<size_t>(edges);
  return e / kHugeChunk + e / kHugeDeg + 2;
}

}  // namespace

size_t tc_workspace_bytes(int nodes, int edges)
{
  const size_t v = static_cast<size_t>(nodes);
  return Carver::round(sizeof(TcControl)) + 7 * Carver::round((v + 1) * 4) + 2 * Carver::round(list_words(edges) * 4) +
         2 * Carver::round(part_words(nodes, edges) * 8) +
         Carver::round(hcnt_words(edges) * 4) + Carver::round(scan_temp_bytes(nodes + 1)) + 4096;
}

int tc_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* launches)
{
  Carver cv(ctx->ws);
  const size_t vp = static_cast<size_t>(nodes) + 1;
  TcControl* ctl = cv.take<TcControl>(1);
  int* ocount = cv.take<int>(vp);
  int* oindex = cv.take<int>(vp);
  int* icount = cv.take<int>(vp);
  int* iindex = cv.take<int>(vp);
  FrontierQueues heavy{nullptr, cv.take<int>(vp), cv.take<int>(vp), cv.take<unsigned>(vp)};
  unsigned long long* mid = cv.take<unsigned long long>(part_words(nodes, edges));
  unsigned long long* big = cv.take<unsigned long long>(part_words(nodes, edges));
  int* olist = cv.take<int>(list_words(edges));
  int* ilist = cv.take<int>(list_words(edges));
  int* hcnt = cv.take<int>(hcnt_words(edges));
  int* temp = cv.take<int>(scan_temp_bytes(nodes + 1) / 4 + 1);
  IG_TRY(zero_async(ctx, ctl, sizeof(TcControl)));
  hipStream_t s = ctx->stream;
  const int gw = grid_for((nodes + kWave - 1) / kWave, kTcBlock / kWave, 1 << 16);
  const int gh = ctx->compute_units * 8;
  const OutSel out{};
  tc_select<false, true><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, out, ocount, oindex, olist, heavy, ctl);
  tc_select_heavy<false, true><<<gh, kTcBlock, 0, s>>>(nindex, nlist, out, ocount, oindex, olist, heavy, hcnt, ctl);
  IG_TRY(exclusive_scan(s, ocount, oindex, nodes, temp));
  tc_select<true, false><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, out, ocount, oindex, olist, heavy, ctl);
  tc_select_heavy<true, false><<<gh, kTcBlock, 0, s>>>(nindex, nlist, out, ocount, oindex, olist, heavy, hcnt, ctl);
  const InSel in{oindex};
  tc_select<false, false><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, in, icount, iindex, ilist, heavy, ctl);
  tc_select_heavy<false, false><<<gh, kTcBlock, 0, s>>>(nindex, nlist, in, icount, iindex, ilist, heavy, hcnt, ctl);
  IG_TRY(exclusive_scan(s, icount, iindex, nodes, temp));
  tc_select<true, false><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, in, icount, iindex, ilist, heavy, ctl);
  tc_select_heavy<true, false><<<gh, kTcBlock, 0, s>>>(nindex, nlist, in, icount, iindex, ilist, heavy, hcnt, ctl);
  tc_count<<<ctx->compute_units * 8, kTcBlock, 0, s>>>(nodes, oindex, olist, iindex, ilist, mid, big, ctl);
  const int hint = ctx->hints[INDIGO_HINT_TC_LDS_LIST];
  const int mid_list = hint > 0 && hint < kMidSlots / 4 ? hint : kMidSlots / 4;
  const int big_list = hint > 0 && hint < kBigSlots / 4 ? hint : kBigSlots / 4;
  IG_TRY(kernel_begin(ctx, "tc_count_group<512, 8192>"));
  tc_count_group<kMidBlock, kMidSlots><<<ctx->compute_units * 4, kMidBlock, 0, s>>>(
      oindex, olist, iindex, ilist, mid, &ctl->nmid, &ctl->next_mid, mid_list, ctl);
  IG_TRY(kernel_end(ctx));
  tc_count_group<kBigBlock, kBigSlots><<<ctx->compute_units, kBigBlock, 0, s>>>(
      oindex, olist, iindex, ilist, big, &ctl->nbig, &ctl->next_big, big_list, ctl);
  IG_HIP_TRY(hipGetLastError());
  if (launches) *launches = 18;
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
