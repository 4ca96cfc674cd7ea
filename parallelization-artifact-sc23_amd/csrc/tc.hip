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
// Keys.  The hubs are the at most 65535 vertices of largest degree (a
// degree threshold: every vertex above it is a hub), and they are numbered
// 0.. in id order; every other vertex v is named by the key 65536 + v.  The
// hubs are closed upwards in the (degree, id) order, so the out-list of a hub
// holds hubs only and is stored as 16-bit keys; on R-MAT graphs three
// quarters of all probes read such lists, which halves their bytes.  Lists
// hold keys, the LDS tables hold keys, and orngk[key] locates a list.  A
// hub's task probes a bitmap of its N+(x) (kHubKeys bits) instead of a hash
// table: one LDS word and a bit extract per key, counted per lane.
//
// Passes (plain launches on the context stream).  Both oriented lists of a
// vertex are subsets of its adjacency list, so each is stored at the
// vertex's own CSR position (olist / ilist[nindex[v] + k], or the 16-bit
// entries of the same bytes for a hub's out-list) and one pass per list
// builds it - no count pass, no scan:
//   0. keys       : degree histogram, hub threshold, hub numbering (a scan
//                  of the hub flags) and (degree, key) per vertex, with the
//                  4-byte (clamped degree, hub number) record the
//                  out-selection gathers; on a graph of at most 65535
//                  vertices every vertex is a hub and its key is its id, in
//                  one pass;
//   1. out-select : N+ lists, order-preserving; lists < kMidDeg by the
//                  in-wave balancer, longer ones binned into mid (one wave
//                  each) and huge (kHugeChunk-entry items, one workgroup
//                  each, compacted per item); also flags unsorted lists and
//                  self loops; records (|N+|, hub number) per vertex and,
//                  per CSR entry, one byte: does the neighbour outrank the
//                  vertex;
//   2. pack       : one workgroup per huge list closes the gaps between its
//                  items;
//   3-4.           the same for the in-lists N-(x): an entry whose byte is
//                  set is out at once, the others read the neighbour's
//                  (|N+|, hub number) in one 4-byte load;
//   5. count, wave pass: a wave grabs up to 64 consecutive vertices at a time
//      through striped counters (a static stride would be badly unbalanced:
//      on R-MAT graphs a vertex's degree follows the bit pattern of its id); a
//      task with 2 <= |N+(x)| <= kWaveList and at most wave_work probes
//      ((|N+(x)| + |N-(x)|) * |N+(x)| bounds them) builds an open-addressing
//      table of N+(x) in LDS and walks the (list, element) pairs 64 per step
//      with the in-wave balancer, one probe each.  Every other task is cut
//      into parts of kPartItems items (out-list, then in-list), queued for
//      pass 6;
//   6. count, workgroup passes: a workgroup grabs a part, hashes N+(x) (or,
//      for a hub, sets its bitmap) and spreads the part's items over its
//      waves (512 threads and a 32 KiB table up to kMidList entries, 1024
//      threads and a 128 KiB table beyond).  Parts bound the work of one grab, so a few heavy vertices
//      cannot form a tail.  A list too long for the table (m > 2^25 edges, or
//      a hint) is hashed a table's worth at a time, and the part's items are
//      probed against each piece in turn (N+(x) is the disjoint union of the
//      pieces, so the counts add).
#include "tc.hpp"

#include <algorithm>
#include <climits>

#include "frontier.hpp"
#include "scan.hpp"

namespace indigo {

namespace {

constexpr int kTcBlock = 256;
constexpr int kWaveList = 256;    // longest N+(x) the wave pass takes
constexpr int kWaveSlots = 1024;  // its largest LDS table (4 slots per entry)
constexpr int kMidBlock = 512;
constexpr int kMidList = 2048;
constexpr int kMidSlots = 8192;
constexpr int kBigBlock = 1024;
constexpr int kBigSlots = 32768;
static_assert(kMidBlock == 512 && kMidSlots == 8192, "the main-kernel name below spells these out");
static_assert(kHugeChunk % kTcBlock == 0, "tc_pack_huge moves whole items");
// Probe bound of a wave-pass task: 2^13 below 2^24 CSR entries, else 2^14.
// The workgroup passes take the rest; on small graphs their parallelism
// beats the wave pass's per-task round trips (R-MAT 14-18: 0.38 -> 0.25,
// 0.71 -> 0.41 and 1.34 -> 0.89 ms against 2^15), while from R-MAT 20 on
// 2^13 puts too many hash tasks there (scale 22: 10.6 -> 11.1 ms) and 2^14
// ties with 2^15.
int wave_work_for(int edges) { return edges < (1 << 24) ? 1 << 13 : 1 << 14; }
constexpr int kPartItems = 2048;    // items per workgroup-pass part (1024 and 4096: slower)
constexpr int kUnroll = 4;          // windows of 64 per step over a long list
constexpr int kLongList = 128;      // lists walked by the whole wave (64, 96: slower; 160, 192: the same)
constexpr int kWalk = 4;            // windows of 64 in flight per step of that walk (2 and 8: slower);
                                    // bitmap tasks, whose windows need fewer registers, take 2 * kWalk
constexpr int kTinyList = 16;       // wave-pass tasks this small are counted kTinyGroup at a time
constexpr int kTinyGroup = 16;      // tables of kTinySlots side by side in the wave's region
constexpr int kTinySlots = 64;
constexpr int kWaveRegion = kWaveSlots;  // a wave's LDS words
static_assert(4 * kTinyList <= kTinySlots && kTinyGroup * kTinySlots <= kWaveRegion && kTinyList < kLongList,
              "tiny-task tables: 64 slots each, all in the wave's region, no long lists");
constexpr int kHubKeys = 1 << 16;  // keys below this number hubs (16-bit rows); others are 65536 + id
constexpr int kMaxHubs = kHubKeys - 1;  // hub numbers 0..65534: 0xffff marks "no hub" in the 4-byte records
constexpr int kRankDeg = (1 << 15) - 1;  // degrees at or above this compare exactly (from nindex)
constexpr int kHistLds = 2048;     // degree bins a workgroup counts in LDS
constexpr int kThreshBlock = 1024;
static_assert(kHubKeys % kThreshBlock == 0, "the threshold pass gives every thread a whole run of bins");

// The wave pass's vertex grabs go through kGrabStripes counters, one per
// 128-byte line: same-address atomics serialise (about 10 ns each), and one
// counter for all 65536 grabs of R-MAT 22 had set the pass's length.
constexpr int kGrabStripes = 16;

struct TcControl {
  FrontierCounters heavy;  // lists of >= kMidDeg entries
  unsigned long long count;
  unsigned nmid, nbig;     // tasks deferred to the workgroup passes
  unsigned next_mid, next_big;
  int error;
  int hub_bin;             // hubs: min(degree, kHubKeys) >= hub_bin
  int pad[24];
  unsigned grab[kGrabStripes][32];  // stripe s: grabs of chunks s, s + kGrabStripes, ... (first word used)
};

IG_DEV bool outranks(int du, int u, int dv, int v) { return du > dv || (du == dv && u > v); }

IG_DEV bool narrow_key(int key) { return key < kHubKeys; }

// Entry k of the row at list + b: 16-bit keys (a hub's out-list, in the first
// half of the row's bytes) or 32-bit keys.
IG_DEV int row_get(const int* list, int b, int k, bool narrow)
{
  return narrow ? static_cast<int>(reinterpret_cast<const unsigned short*>(list + b)[k]) : list[b + k];
}
IG_DEV void row_put(int* list, int b, int k, int key, bool narrow)
{
  if (narrow) reinterpret_cast<unsigned short*>(list + b)[k] = static_cast<unsigned short>(key);
  else list[b + k] = key;
}

// Key of vertex u from the low half of its 4-byte record: its hub number, or
// 0xffff for "not a hub".
IG_DEV int key_of(unsigned rec, int u)
{
  const int h = static_cast<int>(rec & 0xffffu);
  return h != 0xffff ? h : kHubKeys + u;
}

// Out-list of v: the keys of the neighbours that outrank v; 16-bit when v is
// a hub.  A neighbour's rank and key come from one 4-byte record, rk[u] =
// min(degree, kRankDeg) << 16 | hub number (0xffff: none), so the gathers
// read a 16 MB array at scale 22 instead of 32 MB; two clamped degrees
// compare exactly from nindex.  Each CSR entry e of v's row also gets its
// outcome in up[e], so the in-selection needs no rank test.
struct OutSel {
  const int2* __restrict__ dk;  // (degree, key) per vertex
  const unsigned* __restrict__ rk;
  const int* __restrict__ nindex;
  unsigned char* up;            // per CSR entry: 1 when the neighbour outranks the row's vertex
  IG_DEV int own_odeg(int) const { return 0; }
  IG_DEV bool narrow(int v) const { return narrow_key(dk[v].y); }
  IG_DEV bool operator()(int v, int dv, int, int e, int u, int& key) const
  {
    const unsigned r = rk[u];
    key = key_of(r, u);
    const int cu = static_cast<int>(r >> 16), cv = min(dv, kRankDeg);
    bool k;
    if (cu != cv) k = cu > cv;
    else if (cu < kRankDeg) k = u > v;  // equal degrees
    else k = outranks(nindex[u + 1] - nindex[u], u, dv, v);
    up[e] = k ? 1 : 0;
    return k;
  }
};

// In-list of v: the keys of the neighbours u that v outranks (u -> v) whose
// out-list is shorter than v's; always 32-bit.  The out-selection's up[e]
// rules out the neighbours that outrank v (about half the entries) with a
// 1-byte coalesced load; only the others read their 4-byte record
// ok[u] = |N+(u)| << 16 | hub number (|N+| <= sqrt(E) < 2^16).
struct InSel {
  const unsigned char* __restrict__ up;
  const unsigned* __restrict__ ok;
  IG_DEV int own_odeg(int v) const { return static_cast<int>(ok[v] >> 16); }
  IG_DEV bool narrow(int) const { return false; }
  IG_DEV bool operator()(int, int, int odv, int e, int u, int& key) const
  {
    if (up[e] != 0) return false;
    const unsigned r = ok[u];
    key = key_of(r, u);
    return static_cast<int>(r >> 16) < odv;
  }
};

// Where a finished selection is recorded.  Lists live at their vertex's own
// CSR position (list[nindex[v] + k]), so no count pass or scan is needed.
struct OutRec {
  const int2* __restrict__ dk;
  unsigned* ok;  // |N+(v)| << 16 | hub number (0xffff: none) by vertex id
  int2* orng;    // (start, |N+(v)|) by vertex id
  int2* orngk;   // the same by key
  IG_DEV void operator()(int v, int, int beg, int kept) const
  {
    const int key = dk[v].y;
    ok[v] = static_cast<unsigned>(kept) << 16 | static_cast<unsigned>(narrow_key(key) ? key : 0xffff);
    orng[v] = make_int2(beg, kept);
    orngk[key] = make_int2(beg, kept);
  }
};
struct InRec {
  int* ilen;  // |N-(v)|
  IG_DEV void operator()(int v, int, int, int kept) const { ilen[v] = kept; }
};

// Degree histogram over min(degree, kHubKeys): the low bins, where most
// vertices fall, per workgroup in LDS, the rest straight to `hist`.
__global__ __launch_bounds__(kTcBlock) void tc_degree_hist(int nodes, const int* __restrict__ nindex,
                                                           unsigned* hist)
{
  __shared__ unsigned s_h[kHistLds];
  for (int i = threadIdx.x; i < kHistLds; i += kTcBlock) s_h[i] = 0;
  __syncthreads();
  for (int v = blockIdx.x * kTcBlock + threadIdx.x; v < nodes; v += gridDim.x * kTcBlock) {
    const int d = min(nindex[v + 1] - nindex[v], kHubKeys);
    if (d < kHistLds) atomicAdd(&s_h[d], 1u);
    else atomic_add_relaxed(&hist[d], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHistLds; i += kTcBlock)
    if (s_h[i] != 0) atomic_add_relaxed(&hist[i], s_h[i]);
}

// One workgroup: hub_bin = the smallest bin b whose suffix (the vertices in
// bins >= b) holds at most `cap` vertices; kHubKeys + 1 (no hubs) when even
// the top bin holds more.  Suffixes only grow towards lower bins, so each
// thread walks its run of 64 bins downwards until the cap is passed.
__global__ __launch_bounds__(kThreshBlock) void tc_hub_threshold(const unsigned* __restrict__ hist, int cap,
                                                                 TcControl* ctl)
{
  constexpr int kPer = kHubKeys / kThreshBlock;
  static_assert(kPer % 4 == 0, "runs load as uint4");
  constexpr int kWaves = kThreshBlock / kWave;
  __shared__ int s_wave[kWaves];
  __shared__ int s_best;
  const int t = threadIdx.x, lane = lane_id(), wid = t / kWave;
  const int lo = t * kPer;
  // the run in registers, loaded at once (one dependent load per bin had
  // made this kernel 60 us); the top bin joins the last run
  int h[kPer];
#pragma unroll
  for (int q = 0; q < kPer / 4; q++) {
    const uint4 w = reinterpret_cast<const uint4*>(hist + lo)[q];
    h[4 * q] = static_cast<int>(w.x);
    h[4 * q + 1] = static_cast<int>(w.y);
    h[4 * q + 2] = static_cast<int>(w.z);
    h[4 * q + 3] = static_cast<int>(w.w);
  }
  const int top = t == kThreshBlock - 1 ? static_cast<int>(hist[kHubKeys]) : 0;
  int s = top;
#pragma unroll
  for (int q = 0; q < kPer; q++) s += h[q];
  // suffix sums over the lanes: an inclusive scan of the lanes in reverse
  const int inc = wave_inclusive_scan(wave_shfl(s, kWave - 1 - lane));
  const int suf = wave_shfl(inc, kWave - 1 - lane);  // s of this lane and the lanes above
  if (lane == 0) s_wave[wid] = suf;
  if (t == 0) s_best = kHubKeys + 1;
  __syncthreads();
  int acc = suf - s;  // vertices in bins above the run
  for (int w = wid + 1; w < kWaves; w++) acc += s_wave[w];
  int best = kHubKeys + 1;
  acc += top;
  if (t == kThreshBlock - 1 && acc <= cap) best = kHubKeys;
  if (acc <= cap) {
#pragma unroll
    for (int q = kPer - 1; q >= 0; q--) {
      acc += h[q];
      if (acc <= cap) best = lo + q;
    }
  }
  if (best <= kHubKeys) atomicMin(&s_best, best);
  __syncthreads();
  if (t == 0) ctl->hub_bin = s_best;
}

__global__ __launch_bounds__(kTcBlock) void tc_hub_flags(int nodes, const int* __restrict__ nindex,
                                                         const TcControl* ctl, int* flags)
{
  const int hb = ctl->hub_bin;
  for (int v = blockIdx.x * kTcBlock + threadIdx.x; v < nodes; v += gridDim.x * kTcBlock)
    flags[v] = min(nindex[v + 1] - nindex[v], kHubKeys) >= hb ? 1 : 0;
}

// (degree, key) per vertex; hidx: the exclusive scan of the hub flags, or
// null when every vertex is a hub (then a hub's number is its id).
__global__ __launch_bounds__(kTcBlock) void tc_hub_keys(int nodes, const int* __restrict__ nindex,
                                                        const TcControl* ctl, const int* __restrict__ hidx,
                                                        int2* dk, unsigned* rk)
{
  const int hb = hidx ? ctl->hub_bin : 0;
  for (int v = blockIdx.x * kTcBlock + threadIdx.x; v < nodes; v += gridDim.x * kTcBlock) {
    const int d = nindex[v + 1] - nindex[v];
    const int key = !hidx ? v : (min(d, kHubKeys) >= hb ? hidx[v] : kHubKeys + v);
    dk[v] = make_int2(d, key);
    rk[v] = static_cast<unsigned>(min(d, kRankDeg)) << 16 | static_cast<unsigned>(narrow_key(key) ? key : 0xffff);
  }
}

// Selection over short lists (< kMidDeg): a wave takes 64 consecutive
// vertices and walks their CSR entries 64 per step with the in-wave balancer;
// kept entries are compacted per list in entry order (ballot ranks within
// the owner's lanes plus the owner's running count), so selected lists stay
// sorted.  The out pass (kOut) also bins the long lists into the mid / huge
// queues and flags unsorted lists and self loops.
template <bool kOut, typename Sel, typename Rec>
__global__ __launch_bounds__(kTcBlock) IG_SGPR_CAP void tc_select(int nodes, const int* __restrict__ nindex,
                                                      const int* __restrict__ nlist, Sel sel, Rec rec, int* list,
                                                      FrontierQueues heavy, TcControl* ctl)
{
  __shared__ FrontierEmitter s_emit;
  const int lane = lane_id();
  const int wpb = kTcBlock / kWave;
  const int nchunks = (nodes + kWave - 1) / kWave;
  if (kOut) fe_reset(s_emit);
  bool bad = false;
  for (int c = blockIdx.x * wpb + threadIdx.x / kWave; c < nchunks; c += gridDim.x * wpb) {
    const int v = c * kWave + lane;
    const bool valid = v < nodes;
    const int beg = valid ? nindex[v] : 0;
    const int dv = valid ? nindex[v + 1] - beg : 0;
    const bool long_list = dv >= kMidDeg;
    if (kOut) frontier_emit_block(s_emit, long_list, v, dv, heavy, &ctl->heavy);
    const bool mine = valid && !long_list;
    const int odv = mine ? sel.own_odeg(v) : 0;
    const int nar = mine && sel.narrow(v) ? 1 : 0;
    const int slen = mine ? dv : 0;
    const int incl = wave_inclusive_scan(slen);
    const int excl = incl - slen;
    const int total = wave_bcast(incl, kWave - 1);
    int kept = 0;  // as an owner lane: entries of its list kept so far
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      const int obeg = wave_shfl(beg, o);
      const int oexcl = wave_shfl(excl, o);
      const int odeg = wave_shfl(dv, o);
      const int oodv = wave_shfl(odv, o);
      const int okept = wave_shfl(kept, o);
      const bool onar = wave_shfl(nar, o) != 0;
      const int ov = c * kWave + o;
      bool keep = false;
      int key = 0;
      if (k < total) {
        const int e = obeg + (k - oexcl);
        const int u = nlist[e];
        if (kOut) bad |= (u == ov) || (e > obeg && nlist[e - 1] >= u);
        keep = sel(ov, odeg, oodv, e, u, key);
      }
      const unsigned long long m = __ballot(keep);
      // rank among the kept entries of the same owner in this step
      const int seg0 = min(max(oexcl - base, 0), kWave - 1);
      const unsigned long long mine_below = m & ((1ull << lane) - 1ull) & ~((1ull << seg0) - 1ull);
      if (keep) row_put(list, obeg, okept + __popcll(mine_below), key, onar);
      // each owner adds its lanes' kept entries of this step
      const int s0 = min(max(excl - base, 0), kWave), s1 = min(max(incl - base, 0), kWave);
      const int w = s1 - s0;
      const unsigned long long seg = w <= 0 ? 0ull : (w >= kWave ? ~0ull : ((1ull << w) - 1ull) << s0);
      kept += __popcll(m & seg);
    }
    if (mine) rec(v, dv, beg, kept);
  }
  if (kOut) {
    fe_flush(s_emit, heavy, &ctl->heavy);
    if (__ballot(bad) != 0 && lane == 0) st_relaxed(&ctl->error, 1);
  }
}

// Selection over the binned long lists.  Mid lists: one wave each, in order,
// written at the list's own position and recorded.  Huge lists: one
// workgroup per kHugeChunk-entry item, compacted in order at the item's own
// position with the item's count in hcnt; tc_pack_huge then closes the gaps.
template <bool kOut, typename Sel, typename Rec>
__global__ __launch_bounds__(kTcBlock) IG_SGPR_CAP void tc_select_heavy(const int* __restrict__ nindex,
                                                            const int* __restrict__ nlist, Sel sel, Rec rec,
                                                            int* list, FrontierQueues heavy, int* hcnt, TcControl* ctl)
{
  __shared__ int s_owner;
  __shared__ int s_wcnt[kTcBlock / kWave];
  const int lane = lane_id();
  const int wid = threadIdx.x / kWave;
  const unsigned long long hp = ctl->heavy.hpack;
  const unsigned nh = huge_count(hp), nch = huge_chunks(hp);
  bool bad = false;
  for (unsigned item = blockIdx.x; item < nch; item += gridDim.x) {
    if (threadIdx.x == 0) s_owner = huge_owner(heavy.hoff, static_cast<int>(nh), item);
    __syncthreads();
    const int j = s_owner;
    const int v = heavy.hq[j];
    const int vb = nindex[v], ve = nindex[v + 1];
    const int dv = ve - vb;
    const int odv = sel.own_odeg(v);
    const bool nar = sel.narrow(v);
    const int off = static_cast<int>(item - heavy.hoff[j]) * kHugeChunk;  // the item's first entry in the row
    const int beg = vb + off;
    const int end = min(ve, beg + kHugeChunk);
    int run = 0;
    for (int base = beg; base < end; base += kTcBlock) {
      const int e = base + static_cast<int>(threadIdx.x);
      bool keep = false;
      int key = 0;
      if (e < end) {
        const int u = nlist[e];
        if (kOut) bad |= (u == v) || (e > vb && nlist[e - 1] >= u);
        keep = sel(v, dv, odv, e, u, key);
      }
      const unsigned long long m = __ballot(keep);
      if (lane == 0) s_wcnt[wid] = __popcll(m);
      __syncthreads();
      int before = 0, total = 0;
      for (int w = 0; w < kTcBlock / kWave; w++) {
        before += w < wid ? s_wcnt[w] : 0;
        total += s_wcnt[w];
      }
      if (keep) row_put(list, vb, off + run + before + lanes_below(m), key, nar);
      run += total;
      __syncthreads();
    }
    if (threadIdx.x == 0) hcnt[item] = run;
  }
  const unsigned nm = ctl->heavy.nm;
  for (unsigned i = blockIdx.x * (kTcBlock / kWave) + wid; i < nm; i += gridDim.x * (kTcBlock / kWave)) {
    const int v = heavy.mq[i];
    const int vb = nindex[v], ve = nindex[v + 1];
    const int dv = ve - vb;
    const int odv = sel.own_odeg(v);
    const bool nar = sel.narrow(v);
    int k = 0;
    // kUnroll windows per step: their list and neighbour-state loads overlap
    for (int base = vb; base < ve; base += kUnroll * kWave) {
      int u[kUnroll];
#pragma unroll
      for (int q = 0; q < kUnroll; q++) {
        const int e = base + q * kWave + lane;
        u[q] = e < ve ? nlist[e] : -1;
        if (kOut && e < ve) bad |= (u[q] == v) || (e > vb && nlist[e - 1] >= u[q]);
      }
      bool keep[kUnroll];
      int key[kUnroll];
#pragma unroll
      for (int q = 0; q < kUnroll; q++) {
        key[q] = 0;
        keep[q] = u[q] >= 0 && sel(v, dv, odv, base + q * kWave + lane, u[q], key[q]);
      }
#pragma unroll
      for (int q = 0; q < kUnroll; q++) {
        const unsigned long long m = __ballot(keep[q]);
        if (keep[q]) row_put(list, vb, k + lanes_below(m), key[q], nar);
        k += __popcll(m);
      }
    }
    if (lane == 0) rec(v, dv, vb, k);
  }
  if (kOut && __ballot(bad) != 0 && lane == 0) st_relaxed(&ctl->error, 1);
}

// One workgroup per huge list: moves each item's kept entries down to follow
// the previous items' (in item order, so the list stays sorted) and records
// the list.  An item's entries are all read before any is written, and its
// destination never reaches past its own start, so nothing unread is
// overwritten.
template <typename Sel, typename Rec>
__global__ __launch_bounds__(kTcBlock) void tc_pack_huge(const int* __restrict__ nindex, int* list,
                                                         FrontierQueues heavy, const int* __restrict__ hcnt, Sel sel,
                                                         Rec rec, const TcControl* ctl)
{
  constexpr int kPer = kHugeChunk / kTcBlock;
  const unsigned nh = huge_count(ctl->heavy.hpack);
  for (unsigned j = blockIdx.x; j < nh; j += gridDim.x) {
    const int v = heavy.hq[j];
    const int vb = nindex[v], dv = nindex[v + 1] - vb;
    const bool nar = sel.narrow(v);
    const unsigned c0 = heavy.hoff[j];
    const int nc = (dv + kHugeChunk - 1) / kHugeChunk;
    int run = 0;
    for (int c = 0; c < nc; c++) {
      const int k = hcnt[c0 + c];
      int buf[kPer];
#pragma unroll
      for (int q = 0; q < kPer; q++) {
        const int i = q * kTcBlock + static_cast<int>(threadIdx.x);
        buf[q] = i < k ? row_get(list, vb, c * kHugeChunk + i, nar) : 0;
      }
      __syncthreads();
      if (run != c * kHugeChunk) {
#pragma unroll
        for (int q = 0; q < kPer; q++) {
          const int i = q * kTcBlock + static_cast<int>(threadIdx.x);
          if (i < k) row_put(list, vb, run + i, buf[q], nar);
        }
      }
      __syncthreads();
      run += k;
    }
    if (threadIdx.x == 0) rec(v, dv, vb, run);
  }
}

// Not a key and not the empty-slot marker: never found in a table.
constexpr int kNoEntry = -2;

// Hash tables in LDS: buckets of 4 slots (16 bytes, read with one
// ds_read_b128), a power of two >= 8d slots (load <= 1/8) when that fits in
// `cap` slots, else >= 4d (callers keep 4d within cap), and at least 64.
// The home bucket of key w is bits [24 - log2(buckets), 24) of
// (w mod 2^24) * 0x9E3779: Fibonacci hashing in a 24-bit word
// (0x9E3779 / 2^24 ~ 0.618, so consecutive keys land ~0.618 of the table
// apart) with a full-rate 24-bit multiply.  Keys above 2^24 share home
// buckets, which costs probes, never results.  Empty slots hold -1.  An
// insertion takes the first empty slot of its bucket, in slot order, and
// moves on to the next bucket (cyclically) only when all four are taken, so
// the taken slots of a bucket are always a prefix: a bucket whose last slot
// is empty ends every probe sequence through it.  At load 1/8 a bucket is
// full with probability ~0.2%.
struct TcTable {
  int size;       // slots
  unsigned mask;  // buckets - 1
  int shift;      // 24 - log2(buckets)
  int bits;       // log2(buckets)
};

IG_DEV TcTable table_for(int d, int cap)
{
  int size = max(64, 1 << (32 - __clz(8 * d - 1)));
  if (size > cap) size >>= 1;
  const int buckets = size / 4;
  return TcTable{size, static_cast<unsigned>(buckets - 1), __clz(buckets) - 7, 31 - __clz(buckets)};
}

IG_DEV unsigned bucket_of(int w, const TcTable& t)
{
  // one v_bfe_u32 takes the bits
  return __builtin_amdgcn_ubfe(static_cast<unsigned>(__umul24(static_cast<unsigned>(w), 0x9E3779u)),
                               static_cast<unsigned>(t.shift), static_cast<unsigned>(t.bits));
}

IG_DEV int4 read_bucket(const int* tab, unsigned b) { return reinterpret_cast<const int4*>(tab)[b]; }

// (xor-min form: compares of the four elements would be packed into bit
// vectors by the compiler, at three times the instructions)
IG_DEV bool in_bucket(const int4& s, int w)
{
  const unsigned u = static_cast<unsigned>(w);
  return min(min(static_cast<unsigned>(s.x) ^ u, static_cast<unsigned>(s.y) ^ u),
             min(static_cast<unsigned>(s.z) ^ u, static_cast<unsigned>(s.w) ^ u)) == 0u;
}

IG_DEV void table_insert(int* tab, const TcTable& t, int w)
{
  for (unsigned b = bucket_of(w, t);; b = (b + 1) & t.mask) {
    int* s = tab + 4 * b;
    for (int j = 0; j < 4; j++)
      if (atomicCAS(&s[j], -1, w) == -1) return;
  }
}

// Probe from bucket b on, one bucket per LDS read.
IG_DEV bool table_find_from(const int* tab, const TcTable& t, int w, unsigned b)
{
  int4 s = read_bucket(tab, b);
  while (!in_bucket(s, w) && s.w != -1) {
    b = (b + 1) & t.mask;
    s = read_bucket(tab, b);
  }
  return in_bucket(s, w);
}

IG_DEV bool table_find(const int* tab, const TcTable& t, int w) { return table_find_from(tab, t, w, bucket_of(w, t)); }

IG_DEV void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The task of vertex x: N+(x) at row ob of olist (16-bit keys when narrow),
// its in-list at ilist[ib, ib + di).
struct TcTask {
  int ob, d, ib, di;
  bool narrow;
};

// Strided over the calling lanes (i0, step): inserts N+(x)[lo, hi) into tab.
IG_DEV void table_fill(int* tab, const TcTable& t, const int* olist, const TcTask& x, int lo, int hi, int i0,
                       int step)
{
  for (int i = lo + i0; i < hi; i += step) table_insert(tab, t, row_get(olist, x.ob, i, x.narrow));
}

// Hits of a wave: counted on the scalar unit (ballots) by HashProbe, per
// lane by BitProbe, and summed once per count_pairs.
struct Hits {
  unsigned long long wave = 0;  // wave-uniform
  unsigned lane = 0;            // this lane's own
};

// Membership tests of a task's N+(x) in LDS, as policies of the walks below.
// Wave-collective: counts the keys w[0..kN) found.  A key is kNoEntry (never
// found) unless kSure: then every w[q] was read from a 16-bit row, so it is
// below kHubKeys, and the lanes with ok[q] false (past the row's end, holding
// whatever follows it) must not count.
//
// HashProbe: the hash table (any keys): the home buckets of every window at
// once, and the few lanes whose bucket is full and holds no match go on
// through the next ones.  Hits: one ballot per slot, counted per wave.
struct HashProbe {
  static constexpr bool kLaneHits = false;
  const int* tab;
  TcTable t;
  template <int kN, bool kSure>
  IG_DEV void windows(const int* w, const bool*, int lane, Hits& h) const
  {
    static_assert(!kSure, "16-bit rows with unmasked lanes go to BitProbe only");
    unsigned b[kN];
    int4 sl[kN];
#pragma unroll
    for (int q = 0; q < kN; q++) {
      b[q] = bucket_of(w[q], t);
      sl[q] = read_bucket(tab, b[q]);
    }
    // hit masks straight from one compare per slot, on the scalar unit
#pragma unroll
    for (int q = 0; q < kN; q++) {
      unsigned long long hm = __ballot(sl[q].x == w[q]) | __ballot(sl[q].y == w[q]) | __ballot(sl[q].z == w[q]) |
                              __ballot(sl[q].w == w[q]);
      const unsigned long long go = __ballot(sl[q].w != -1) & ~hm;  // full bucket, no match
      if (go != 0) {
        const bool mine = ((go >> lane) & 1ull) != 0;
        hm |= __ballot(mine && table_find_from(tab, t, w[q], (b[q] + 1) & t.mask));
      }
      h.wave += static_cast<unsigned long long>(__popcll(hm));
    }
  }
};

// BitProbe: a bitmap of hub keys (kHubKeys bits, then one zero word), for a
// hub's task: its N+(x) holds hubs only.  A key k reads word
// min(k, kHubKeys) / 32, so every other key (65536 + id, and kNoEntry as
// unsigned) reads the zero word; the hit is bit k mod 32, added to the
// lane's own count.  Per key: one 4-byte LDS read and four VALU
// instructions when kSure (address shift and mask, bit extract, add), against
// a 16-byte bucket read, a hash and four compares.
struct BitProbe {
  static constexpr bool kLaneHits = true;
  static constexpr int kWords = kHubKeys / 32 + 1;
  static constexpr int kHistAt = 4096;  // the waves' owner histograms follow the bitmap in LDS
  const unsigned* bits;
  int* hist;  // this wave's 2 * kWalk * 64 ints (16-byte aligned), for balanced_step
  template <int kN, bool kSure>
  IG_DEV void windows(const int* w, const bool* ok, int, Hits& h) const
  {
    unsigned v[kN];
#pragma unroll
    for (int q = 0; q < kN; q++) {
      const unsigned k = static_cast<unsigned>(w[q]);
      v[q] = bits[(kSure ? k : min(k, static_cast<unsigned>(kHubKeys))) >> 5];
    }
#pragma unroll
    for (int q = 0; q < kN; q++) {  // v_bfe_u32 takes the offset mod 32 by itself
      const unsigned hit = __builtin_amdgcn_ubfe(v[q], static_cast<unsigned>(w[q]), 1u);
      h.lane += !kSure || ok[q] ? hit : 0u;
    }
  }
};

// Wave-collective: kN windows of 64 entries of row[0, ol) from `base`
// (windows past base + 64*kN are not touched), then counted.  Every lane
// loads, from one address with the windows at immediate offsets, so the
// loads issue back to back; unless kFull (all kN windows inside the row)
// the lanes past the row's end read what follows it, which the list arrays'
// padding keeps inside their allocation (kListPad).  Those lanes are masked:
// 16-bit rows under BitProbe keep the value (a valid bitmap index) and mask
// the lane's count, the others take kNoEntry.
template <typename T, int kN, bool kFull, typename P>
IG_DEV void walk_step(const T* row, int base, int ol, int lane, const P& pr, Hits& h)
{
  static_assert(kN >= 1 && kN <= 2 * kWalk, "window count");
  constexpr bool kSure = sizeof(T) == 2 && P::kLaneHits;
  int w[kN];
  bool ok[kN];
  const T* at = row + base + lane;
#pragma unroll
  for (int q = 0; q < kN; q++) {
    const int v = at[q * kWave];
    ok[q] = kFull || base + q * kWave + lane < ol;
    w[q] = kFull || kSure || ok[q] ? v : kNoEntry;
  }
  pr.template windows<kN, kSure>(w, ok, lane, h);
}

// Wave-collective: the whole of row[0, ol): steps of kWalk whole windows
// (2 * kWalk under lane-counted probes, while that many are left), then one
// step of the 1 to kWalk windows left (the last of them partial).
template <typename T, typename P>
IG_DEV void walk_row(const T* row, int ol, int lane, const P& pr, Hits& h)
{
  static_assert(kWalk == 4, "the switch below covers 1 to kWalk windows");
  int base = 0;
  if constexpr (P::kLaneHits) {
    for (; ol - base >= 2 * kWalk * kWave; base += 2 * kWalk * kWave)
      walk_step<T, 2 * kWalk, true>(row, base, ol, lane, pr, h);
  }
  for (; ol - base >= kWalk * kWave; base += kWalk * kWave) walk_step<T, kWalk, true>(row, base, ol, lane, pr, h);
  switch ((ol - base + kWave - 1) / kWave) {
    case 1: walk_step<T, 1, false>(row, base, ol, lane, pr, h); break;
    case 2: walk_step<T, 2, false>(row, base, ol, lane, pr, h); break;
    case 3: walk_step<T, 3, false>(row, base, ol, lane, pr, h); break;
    case 4: walk_step<T, 4, false>(row, base, ol, lane, pr, h); break;
    default: break;
  }
}

// Wave-collective: kN windows of 64 pairs of the in-wave balancer from
// `base` (< total): pair k of the lanes' concatenated short rows is
// list[shift[o] + k] for its owner lane o = #{lanes i : incl_i <= k} (incl:
// the inclusive prefix of the row lengths).  Owners first, then every lane
// loads (past the end: window 0's first pair again) so that the loads issue
// back to back, then counted.
//
// Owners: with an LDS histogram (policies that carry one) the rows ending
// inside the step's windows add 1 at their end position, and one wave scan
// per window gives every lane its owner (about 10 VALU instructions a
// window); else a binary search over the lanes' incl through 6 shuffles
// (about 45 a window).
template <typename T, int kN, typename P>
IG_DEV void balanced_step(const T* __restrict__ list, int incl, unsigned shift, int total, int base, int lane,
                          const P& pr, Hits& h)
{
  static_assert(kN >= 1 && kN <= 2 * kWalk, "window count");
  unsigned at[kN];
  if constexpr (P::kLaneHits) {
    int* hs = pr.hist;
    reinterpret_cast<int4*>(hs)[lane] = make_int4(0, 0, 0, 0);  // kWalk * 64 ints, twice for 2 * kWalk windows
    if (kN > kWalk) reinterpret_cast<int4*>(hs)[kWave + lane] = make_int4(0, 0, 0, 0);
    wave_lds_sync();
    const int qe = incl - base;  // this lane's row ends before pair base + qe
    if (qe >= 1 && qe < kN * kWave) atomicAdd(&hs[qe], 1);
    wave_lds_sync();
    int carry = __popcll(__ballot(qe <= 0));  // rows ended by pair `base`
#pragma unroll
    for (int q = 0; q < kN; q++) {
      const int sc = wave_inclusive_scan(hs[q * kWave + lane]);
      const int o = min(carry + sc, kWave - 1);
      carry += wave_bcast(sc, kWave - 1);
      at[q] = wave_shfl(shift, o) + static_cast<unsigned>(base + q * kWave + lane);
    }
    wave_lds_sync();  // the reads above before the next step's clear
  } else {
#pragma unroll
    for (int q = 0; q < kN; q++) {
      const int k = base + q * kWave + lane;
      const int o = min(wave_owner(incl, k), kWave - 1);
      at[q] = wave_shfl(shift, o) + static_cast<unsigned>(k);
    }
  }
  const unsigned fb = static_cast<unsigned>(wave_bcast(static_cast<int>(at[0]), 0));  // pair `base`, which exists
  int w[kN];
#pragma unroll
  for (int q = 0; q < kN; q++) {
    const bool ok = base + q * kWave + lane < total;
    const int v = list[ok ? at[q] : fb];
    w[q] = ok ? v : kNoEntry;
  }
  pr.template windows<kN, false>(w, nullptr, lane, h);
}

// Wave-collective: each lane's short row of slen entries from list[start]
// (in entries of T), all through the in-wave balancer: kWalk windows of
// pairs per step (2 * kWalk under lane-counted probes, while that many are
// left), then one step of just the windows left.
template <typename T, typename P>
IG_DEV void balanced_rows(const T* __restrict__ list, unsigned start, int slen, int lane, const P& pr, Hits& h)
{
  const int incl = wave_inclusive_scan(slen);
  const int total = wave_bcast(incl, kWave - 1);
  if (total == 0) return;
  const unsigned shift = start - static_cast<unsigned>(incl - slen);  // pair k is list[shift[o] + k]
  int base = 0;
  if constexpr (P::kLaneHits) {
    for (; total - base >= 2 * kWalk * kWave; base += 2 * kWalk * kWave)
      balanced_step<T, 2 * kWalk>(list, incl, shift, total, base, lane, pr, h);
  }
  for (; total - base > (kWalk - 1) * kWave; base += kWalk * kWave)
    balanced_step<T, kWalk>(list, incl, shift, total, base, lane, pr, h);
  switch ((total - base + kWave - 1) / kWave) {
    case 1: balanced_step<T, 1>(list, incl, shift, total, base, lane, pr, h); break;
    case 2: balanced_step<T, 2>(list, incl, shift, total, base, lane, pr, h); break;
    case 3: balanced_step<T, 3>(list, incl, shift, total, base, lane, pr, h); break;
    default: break;
  }
}

// Wave-collective: over the item keys items[ub0 + 64k + lane] (k = 0, 1, ...
// stepping by ustep) of the sequence items_a[0, n_a) (16-bit keys when
// a_narrow), then items_b[0, n_b), so that one batch of 64 can span both
// (a task's N+(x) and N-(x) in one chain of round trips), counts the keys in
// N+(item) found in the table, for the items whose list has at most `limit`
// entries.  Lists of kLongList or more entries are walked by the whole wave,
// one at a time, kWalk (bitmap probes: 2 * kWalk) windows of 64 per step so
// that their loads overlap; shorter ones share windows of 64 pairs through
// the in-wave balancer, as many windows per step, the 16-bit rows and the
// 32-bit rows in turn.  Returns the count in lane 0.
template <typename P>
IG_DEV unsigned long long count_pairs(const int2* __restrict__ orngk, const int* __restrict__ olist,
                                      const void* items_a, bool a_narrow, int n_a, const int* __restrict__ items_b,
                                      int n_b, int limit, int ub0, int ustep, const P& pr)
{
  const int lane = lane_id();
  const unsigned short* olist16 = reinterpret_cast<const unsigned short*>(olist);
  const int n = n_a + n_b;
  Hits h;
  for (int ub = ub0; ub < n; ub += ustep) {
    const int i = ub + lane;
    int key = -1;
    if (i < n_a)
      key = a_narrow ? static_cast<int>(static_cast<const unsigned short*>(items_a)[i])
                     : static_cast<const int*>(items_a)[i];
    else if (i < n)
      key = items_b[i - n_a];
    const int2 r = key >= 0 ? orngk[key] : make_int2(0, 0);
    const bool nar = narrow_key(key);
    const int len = r.y <= limit ? r.y : 0;
    unsigned long long todo = __ballot(len >= kLongList);
    while (todo != 0) {
      const int o = __ffsll(static_cast<long long>(todo)) - 1;
      todo &= todo - 1;
      const int ol = wave_bcast(len, o);
      const int b = wave_bcast(r.x, o);
      if (wave_bcast(nar ? 1 : 0, o) != 0) walk_row(reinterpret_cast<const unsigned short*>(olist + b), ol, lane, pr, h);
      else walk_row(olist + b, ol, lane, pr, h);
    }
    const bool sh = len < kLongList;
    balanced_rows(olist16, 2u * static_cast<unsigned>(r.x), sh && nar ? len : 0, lane, pr, h);
    balanced_rows(olist, static_cast<unsigned>(r.x), sh && !nar ? len : 0, lane, pr, h);
  }
  unsigned long long c = h.wave;
  if (P::kLaneHits) c += wave_sum(h.lane);  // at most a part's probes (< 2^32)
  return lane == 0 ? c : 0ull;
}

// Items [lo, hi) of the task's item sequence (N+(x), then N-(x)).
template <typename P>
IG_DEV unsigned long long task_pairs(const int2* __restrict__ orngk, const int* __restrict__ olist,
                                     const int* __restrict__ ilist, const TcTask& x, int lo, int hi, int ub0,
                                     int ustep, const P& pr)
{
  const int a0 = min(lo, x.d), a1 = min(hi, x.d);
  const int b0 = max(lo, x.d) - x.d, b1 = max(hi, x.d) - x.d;
  const void* items_a = x.narrow ? static_cast<const void*>(reinterpret_cast<const unsigned short*>(olist + x.ob) + a0)
                                 : static_cast<const void*>(olist + x.ob + a0);
  return count_pairs(orngk, olist, items_a, x.narrow, a1 - a0, ilist + x.ib + b0, b1 - b0, x.d, ub0, ustep, pr);
}

// Wave-collective: the tiny tasks (|N+(x)| <= kTinyList) of the lanes in
// `mask` (at most kTinyGroup of them), counted together: each gets a 64-slot
// table in the wave's LDS region, and their items (N+(x), then N-(x)) and
// then the items' list entries are spread over the wave with the in-wave
// balancer.  A counted item's list is no longer than its task's N+(x) (the
// shorter-list rule), so at most kTinyList entries: no long-list walk.  One
// task at a time would pay its own chain of round trips for a handful of
// probes.  Rows of either width mix here, so every entry is read with its
// own row's width.
IG_DEV unsigned long long tiny_tasks(unsigned long long mask, const int2 rl, int dl, int il, int nl,
                                     const int2* __restrict__ orngk, const int* __restrict__ olist,
                                     const int* __restrict__ ilist, int* tab)
{
  const int lane = lane_id();
  // lane j takes the task of the j-th lane of `mask`; lanes past the last
  // task hold an empty one
  int src = kWave;
  {
    unsigned long long m = mask;
    for (int j = 0; m != 0; j++) {
      const int sl = __ffsll(static_cast<long long>(m)) - 1;
      m &= m - 1;
      if (lane == j) src = sl;
    }
  }
  const int ntask = __popcll(mask);
  const bool has = src < kWave;
  const int tob = wave_shfl(rl.x, has ? src : 0);  // in-lists share the out-lists' positions
  const int td0 = wave_shfl(dl, has ? src : 0);
  const int ti0 = wave_shfl(il, has ? src : 0);
  const int tnar = wave_shfl(nl, has ? src : 0);
  const int td = has ? td0 : 0;
  const int tn = has ? td0 + ti0 : 0;  // items: N+(x), then N-(x)
  const TcTable t = table_for(kTinyList, kTinySlots);  // 64 slots for every task
  const int stride = t.size;
  for (int i = lane; i < ntask * stride; i += kWave) tab[i] = -1;
  wave_lds_sync();
  {
    const int incl = wave_inclusive_scan(td);
    const int total = wave_bcast(incl, kWave - 1);
    for (int base = 0; base < total; base += kWave) {
      const int k = base + lane;
      const int j = min(wave_owner(incl, k), kWave - 1);
      const int i = k - wave_shfl(incl - td, j);
      const int jb = wave_shfl(tob, j);
      const bool jn = wave_shfl(tnar, j) != 0;
      if (k < total) table_insert(tab + j * stride, t, row_get(olist, jb, i, jn));
    }
  }
  wave_lds_sync();
  unsigned long long cnt = 0;
  const int incl_n = wave_inclusive_scan(tn);
  const int total_n = wave_bcast(incl_n, kWave - 1);
  for (int ibase = 0; ibase < total_n; ibase += kWave) {
    const int k = ibase + lane;
    const int j = min(wave_owner(incl_n, k), kWave - 1);
    const int i = k - wave_shfl(incl_n - tn, j);
    const int jd = wave_shfl(td, j);
    const int job = wave_shfl(tob, j);
    const bool jn = wave_shfl(tnar, j) != 0;
    int b = 0, len = 0, nk = 0;
    if (k < total_n) {
      const int u = i < jd ? row_get(olist, job, i, jn) : ilist[job + (i - jd)];
      const int2 r = orngk[u];
      b = r.x;
      len = r.y <= jd ? r.y : 0;  // the shorter-list rule of count_pairs
      nk = narrow_key(u) ? 1 : 0;
    }
    const int incl = wave_inclusive_scan(len);
    const int total = wave_bcast(incl, kWave - 1);
    const int excl = incl - len;
    for (int pbase = 0; pbase < total; pbase += kWave) {
      const int kk = pbase + lane;
      const int o = min(wave_owner(incl, kk), kWave - 1);
      const int ob = wave_shfl(b, o);
      const int oi = kk - wave_shfl(excl, o);
      const bool on = wave_shfl(nk, o) != 0;
      const int tj = wave_shfl(j, o);
      if (kk < total) cnt += table_find(tab + tj * stride, t, row_get(olist, ob, oi, on)) ? 1ull : 0ull;
    }
  }
  wave_lds_sync();
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

__global__ __launch_bounds__(kTcBlock) IG_SGPR_CAP void tc_count(int nodes, int grab, int wave_work,
                                                     const int2* __restrict__ dk,
                                                     const int2* __restrict__ orng, const int2* __restrict__ orngk,
                                                     const int* __restrict__ olist, const int* __restrict__ ilen,
                                                     const int* __restrict__ ilist, unsigned long long* mid,
                                                     unsigned long long* big, TcControl* ctl)
{
  __shared__ __align__(16) int s_tab[kTcBlock / kWave][kWaveRegion];
  const int lane = lane_id();
  int* tab = s_tab[threadIdx.x / kWave];
  unsigned long long cnt = 0;
  // chunks of `grab` (at most 64) consecutive vertices, one per lane: 64 on
  // large graphs (8 grabs per wave on average at R-MAT 22: coarser grabs
  // leave a tail), fewer on small ones, whose few grabs would otherwise
  // leave one wave with a long serial run of tasks.  Chunk c is grabbed
  // through stripe c mod kGrabStripes; a wave starts at its own stripe and
  // moves on when it runs dry, and leaves after finding every stripe dry in
  // a row (a dry stripe stays dry).
  const unsigned nch = static_cast<unsigned>((nodes + grab - 1) / grab);
  int stripe = static_cast<int>((blockIdx.x * (kTcBlock / kWave) + threadIdx.x / kWave) % kGrabStripes);
  int dry = 0;
  for (;;) {
    unsigned c = 0;
    if (lane == 0) {
      // a dry stripe is seen with a load, without another atomic
      unsigned* g = &ctl->grab[stripe][0];
      const unsigned k = ld_relaxed(g) * kGrabStripes + static_cast<unsigned>(stripe) < nch ? atomic_add_relaxed(g, 1u)
                                                                                             : ld_relaxed(g);
      c = k * kGrabStripes + static_cast<unsigned>(stripe);
    }
    c = wave_shfl(c, 0);
    if (c >= nch) {
      if (++dry == kGrabStripes) break;
      stripe = (stripe + 1) % kGrabStripes;
      continue;
    }
    dry = 0;
    const unsigned v0 = c * static_cast<unsigned>(grab);
    // heavy tasks go to the workgroup passes in parts (|N+(x)| < 2: no
    // triangle can be found through N+(x))
    const int vl = static_cast<int>(v0) + lane;
    const bool mine = lane < grab && vl < nodes;
    const int2 rl = mine ? orng[vl] : make_int2(0, 0);
    const int nl = mine && narrow_key(dk[vl].y) ? 1 : 0;
    const int dl = rl.y;
    const int il = mine ? ilen[vl] : 0;
    const bool here = dl >= 2 && dl <= kWaveList && (dl + il) * dl <= wave_work;
    const int parts = dl >= 2 && !here ? (dl + il + kPartItems - 1) / kPartItems : 0;
    append_parts(dl <= kMidList ? parts : 0, vl, mid, &ctl->nmid);
    append_parts(dl > kMidList ? parts : 0, vl, big, &ctl->nbig);
    unsigned long long todo = __ballot(here);
    // tiny tasks first, kTinyGroup at a time
    unsigned long long tiny = __ballot(here && dl <= kTinyList);
    todo &= ~tiny;
    while (tiny != 0) {
      unsigned long long grp = 0;
      for (int q = 0; q < kTinyGroup && tiny != 0; q++) {
        const unsigned long long low = tiny & (~tiny + 1ull);
        grp |= low;
        tiny &= tiny - 1;
      }
      cnt += tiny_tasks(grp, rl, dl, il, nl, orngk, olist, ilist, tab);
    }
    while (todo != 0) {
      const int j = __ffsll(static_cast<long long>(todo)) - 1;
      todo &= todo - 1;
      const int ob = wave_shfl(rl.x, j);  // in-lists share the out-lists' positions
      const TcTask x{ob, wave_shfl(dl, j), ob, wave_shfl(il, j), wave_shfl(nl, j) != 0};
      const TcTable t = table_for(x.d, kWaveSlots);
      for (int i = lane; i < t.size; i += kWave) tab[i] = -1;
      wave_lds_sync();
      table_fill(tab, t, olist, x, 0, x.d, lane, kWave);
      wave_lds_sync();
      cnt += task_pairs(orngk, olist, ilist, x, 0, x.d + x.di, 0, kWave, HashProbe{tab, t});
      wave_lds_sync();
    }
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt != 0) atomic_add_relaxed(&ctl->count, cnt);
}

// One workgroup per part from `parts` (grabbed through *next).  A hub's
// N+(x) becomes a bitmap of hub keys in LDS (BitProbe); any other N+(x) is
// hashed into LDS lds_list (<= kSlots / 4) entries at a time - one piece
// unless the list is longer - and the part's items are probed against every
// piece.
template <int kBlock, int kSlots>
__global__ __launch_bounds__(kBlock) IG_SGPR_CAP void tc_count_group(const int2* __restrict__ dk, const int2* __restrict__ orng,
                                                         const int2* __restrict__ orngk,
                                                         const int* __restrict__ olist,
                                                         const int* __restrict__ ilen,
                                                         const int* __restrict__ ilist,
                                                         const unsigned long long* __restrict__ parts,
                                                         const unsigned* nparts, unsigned* next, int lds_list,
                                                         TcControl* ctl)
{
  __shared__ __align__(16) int s_tab[kSlots];
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
    const int2 r = orng[v];
    const TcTask x{r.x, r.y, r.x, ilen[v], narrow_key(dk[v].y)};
    const int hi = min(lo + kPartItems, x.d + x.di);
    if (x.narrow) {
      // a hub's N+(x) as a bitmap of hub keys (kHubKeys bits, then a zero
      // word)
      static_assert(BitProbe::kWords <= BitProbe::kHistAt && BitProbe::kHistAt + kWaves * 2 * kWalk * kWave <= kSlots,
                    "the bitmap and the owner histograms fit the table's LDS");
      unsigned* bits = reinterpret_cast<unsigned*>(s_tab);
      for (int i = threadIdx.x; i < BitProbe::kWords; i += kBlock) bits[i] = 0u;
      __syncthreads();
      for (int i = threadIdx.x; i < x.d; i += kBlock) {
        const unsigned k = static_cast<unsigned>(row_get(olist, x.ob, i, true));
        atomicOr(&bits[k >> 5], 1u << (k & 31u));
      }
      __syncthreads();
      cnt += task_pairs(orngk, olist, ilist, x, lo, hi, wid * kWave, kWaves * kWave,
                        BitProbe{bits, s_tab + BitProbe::kHistAt + wid * 2 * kWalk * kWave});
      __syncthreads();
    } else {
      // N+(x) hashed lds_list entries at a time: one piece unless the list
      // is longer; every piece is probed with all of the part's items
      for (int c0 = 0; c0 < x.d; c0 += lds_list) {
        const int c1 = min(x.d, c0 + lds_list);
        const TcTable t = table_for(c1 - c0, kSlots);
        for (int i = threadIdx.x; i < t.size; i += kBlock) s_tab[i] = -1;
        __syncthreads();
        table_fill(s_tab, t, olist, x, c0, c1, threadIdx.x, kBlock);
        __syncthreads();
        cnt += task_pairs(orngk, olist, ilist, x, lo, hi, wid * kWave, kWaves * kWave, HashProbe{s_tab, t});
        __syncthreads();
      }
    }
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0 && cnt != 0) atomic_add_relaxed(&ctl->count, cnt);
}

int grid_for(long long items, int per_block, int cap)
{
  const long long g = (items + per_block - 1) / per_block;
  return static_cast<int>(g < 1 ? 1 : (g > cap ? cap : g));
}

// The walks read up to kWalk * 64 - 1 entries past a row's end (see
// walk_step): at most that many 32-bit words past the last row.
constexpr size_t kListPad = kWalk * kWave + 16;
size_t list_words(int edges) { return static_cast<size_t>(edges) + kListPad; }

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
  // the takes of tc_run, in its order
  const size_t v = static_cast<size_t>(nodes), vp = v + 1;
  return Carver::round(sizeof(TcControl)) + 2 * Carver::round(v * 4) + Carver::round(v * sizeof(int2)) +
         Carver::round((kHubKeys + v) * sizeof(int2)) + Carver::round(v * sizeof(int2)) + Carver::round(v * 4) +
         3 * Carver::round(vp * 4) + 2 * Carver::round(part_words(nodes, edges) * 8) +
         2 * Carver::round(list_words(edges) * 4) + Carver::round(hcnt_words(edges) * 4) +
         Carver::round((kHubKeys + 1) * 4) + Carver::round(v * 4) + Carver::round(vp * 4) +
         Carver::round(scan_temp_bytes(nodes)) + Carver::round(static_cast<size_t>(edges)) + 4096;
}

int tc_run(indigo_ctx* ctx, int nodes, int edges, const int* nindex, const int* nlist, int* launches)
{
  if (nodes > INT_MAX - kHubKeys) return set_error(INDIGO_ERR_INVALID_ARG, "tc: too many vertices for 32-bit keys");
  Carver cv(ctx->ws);
  const size_t vn = static_cast<size_t>(nodes), vp = vn + 1;
  TcControl* ctl = cv.take<TcControl>(1);
  unsigned* ok = cv.take<unsigned>(vn);
  unsigned* rk = cv.take<unsigned>(vn);
  int2* orng = cv.take<int2>(vn);
  int2* orngk = cv.take<int2>(kHubKeys + vn);
  int2* dk = cv.take<int2>(vn);
  int* ilen = cv.take<int>(vn);
  FrontierQueues heavy{nullptr, cv.take<int>(vp), cv.take<int>(vp), cv.take<unsigned>(vp)};
  unsigned long long* mid = cv.take<unsigned long long>(part_words(nodes, edges));
  unsigned long long* big = cv.take<unsigned long long>(part_words(nodes, edges));
  int* olist = cv.take<int>(list_words(edges));
  int* ilist = cv.take<int>(list_words(edges));
  int* hcnt = cv.take<int>(hcnt_words(edges));
  unsigned* hist = cv.take<unsigned>(kHubKeys + 1);
  int* flags = cv.take<int>(vn);
  int* hidx = cv.take<int>(vp);
  int* stemp = cv.take<int>(scan_temp_bytes(nodes) / 4 + 1);
  unsigned char* up = cv.take<unsigned char>(static_cast<size_t>(edges));
  IG_TRY(zero_async(ctx, ctl, sizeof(TcControl)));
  hipStream_t s = ctx->stream;
  const int gw = share_grid(ctx, grid_for((nodes + kWave - 1) / kWave, kTcBlock / kWave, 1 << 16), kTcBlock);
  const int gh = share_grid(ctx, ctx->compute_units * 8, kTcBlock);
  const int gv = share_grid(ctx, grid_for(nodes, kTcBlock, ctx->compute_units * 8), kTcBlock);
  int n_launch = 11;
  // keys: every vertex is a hub (key = id) when they all fit, else the
  // highest-degree ones up to the cap, numbered in id order
  const int hub_hint = ctx->hints[INDIGO_HINT_TC_HUBS];
  const int hub_cap = hub_hint > 0 && hub_hint < kMaxHubs ? hub_hint : kMaxHubs;
  if (nodes <= hub_cap) {
    tc_hub_keys<<<gv, kTcBlock, 0, s>>>(nodes, nindex, ctl, nullptr, dk, rk);
  } else {
    IG_TRY(zero_async(ctx, hist, (kHubKeys + 1) * sizeof(unsigned)));
    // one workgroup per CU: each flushes its busy low bins with global
    // atomics on the same addresses as every other workgroup's
    tc_degree_hist<<<std::min(gv, share_grid(ctx, ctx->compute_units, kTcBlock)), kTcBlock, 0, s>>>(nodes, nindex,
                                                                                                  hist);
    tc_hub_threshold<<<1, kThreshBlock, 0, s>>>(hist, hub_cap, ctl);
    tc_hub_flags<<<gv, kTcBlock, 0, s>>>(nodes, nindex, ctl, flags);
    IG_TRY(exclusive_scan(s, flags, hidx, nodes, stemp));
    tc_hub_keys<<<gv, kTcBlock, 0, s>>>(nodes, nindex, ctl, hidx, dk, rk);
    n_launch += 7;
  }
  const OutSel out{dk, rk, nindex, up};
  const OutRec orec{dk, ok, orng, orngk};
  tc_select<true><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, out, orec, olist, heavy, ctl);
  tc_select_heavy<true><<<gh, kTcBlock, 0, s>>>(nindex, nlist, out, orec, olist, heavy, hcnt, ctl);
  tc_pack_huge<<<gh, kTcBlock, 0, s>>>(nindex, olist, heavy, hcnt, out, orec, ctl);
  const InSel in{up, ok};
  const InRec irec{ilen};
  tc_select<false><<<gw, kTcBlock, 0, s>>>(nodes, nindex, nlist, in, irec, ilist, heavy, ctl);
  tc_select_heavy<false><<<gh, kTcBlock, 0, s>>>(nindex, nlist, in, irec, ilist, heavy, hcnt, ctl);
  tc_pack_huge<<<gh, kTcBlock, 0, s>>>(nindex, ilist, heavy, hcnt, in, irec, ctl);
  {
    // vertices per grab: 64 from 2^18 vertices on, else nodes / 4096 (at
    // least 1), so a small graph still spreads over 4096 grabs while the
    // grabs' same-address atomics stay few
    const int grab = std::max(1, std::min(kWave, nodes / 4096));
    tc_count<<<share_grid(ctx, ctx->compute_units * 8, kTcBlock), kTcBlock, 0, s>>>(
        nodes, grab, wave_work_for(edges), dk, orng, orngk, olist, ilen, ilist, mid, big, ctl);
  }
  const int hint = ctx->hints[INDIGO_HINT_TC_LDS_LIST];
  const int mid_list = hint > 0 && hint < kMidSlots / 4 ? hint : kMidSlots / 4;
  const int big_list = hint > 0 && hint < kBigSlots / 4 ? hint : kBigSlots / 4;
  IG_TRY(kernel_begin(ctx, "tc_count_group<512, 8192>"));
  tc_count_group<kMidBlock, kMidSlots><<<share_grid(ctx, ctx->compute_units * 4, kMidBlock), kMidBlock, 0, s>>>(
      dk, orng, orngk, olist, ilen, ilist, mid, &ctl->nmid, &ctl->next_mid, mid_list, ctl);
  IG_TRY(kernel_end(ctx));
  tc_count_group<kBigBlock, kBigSlots><<<share_grid(ctx, ctx->compute_units, kBigBlock), kBigBlock, 0, s>>>(
      dk, orng, orngk, olist, ilen, ilist, big, &ctl->nbig, &ctl->next_big, big_list, ctl);
  IG_HIP_TRY(hipGetLastError());
  if (launches) *launches = n_launch;
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
