// Host-side CSR construction, synthetic generators and .egr I/O.
// See include/indigo_graph.h for the contract.  Host C++17 only.
#include "indigo_graph.h"
#include "indigo_hip.h"
#include "errors.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

namespace {

using indigo::set_error;

inline uint64_t mix64(uint64_t x)
{
  // splitmix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
  // number generators", OOPSLA 2014)
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Counter-based stream: draw k of stream s is mix64(key(s) + k * golden).
struct Stream {
  uint64_t state;
  explicit Stream(uint64_t seed, uint64_t id) : state(mix64(seed ^ mix64(id + 0x632BE59BD9B4E019ull))) {}
  uint64_t next() { state += 0x9E3779B97F4A7C15ull; return mix64(state); }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

inline int edge_weight(uint64_t seed, int u, int v, int wmin, int wmax)
{
  const uint64_t lo = static_cast<uint32_t>(std::min(u, v));
  const uint64_t hi = static_cast<uint32_t>(std::max(u, v));
  const uint64_t h = mix64(seed ^ mix64((lo << 32) | hi));
  const uint64_t range = static_cast<uint64_t>(wmax - wmin) + 1;
  return wmin + static_cast<int>(h % range);
}

int pick_threads(int threads)
{
  if (threads > 0) return std::min(threads, 64);
  const unsigned hc = std::thread::hardware_concurrency();
  return static_cast<int>(std::max(1u, std::min(hc, 16u)));
}

template <typename F>
void parallel_for(long long n, int threads, F&& fn)
{
  if (n <= 0) return;
  const int t = static_cast<int>(std::min<long long>(threads, std::max<long long>(1, n / 4096 + 1)));
  if (t <= 1) {
    fn(0LL, n, 0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t);
  for (int i = 0; i < t; i++) {
    const long long b = n * i / t, e = n * (i + 1) / t;
    pool.emplace_back([&fn, b, e, i] { fn(b, e, i); });
  }
  for (auto& th : pool) th.join();
}

// fn(i) on k threads, i = 0..k-1.
template <typename F>
void run_threads(int k, F&& fn)
{
  if (k <= 1) {
    fn(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(k);
  for (int i = 0; i < k; i++) pool.emplace_back([&fn, i] { fn(i); });
  for (auto& th : pool) th.join();
}

bool weights_wanted(int wmin, int wmax) { return wmin >= 1 && wmax >= wmin; }

void clear(indigo_host_graph* g)
{
  g->nodes = 0;
  g->edges = 0;
  g->nindex = nullptr;
  g->nlist = nullptr;
  g->eweight = nullptr;
}

// Edge-list -> simple undirected CSR.  Counting sort by source, then each
// list sorted and de-duplicated in parallel, then compacted.
int build_csr_impl(int nodes, long long npairs, const int* src, const int* dst, bool symmetrize,
                   uint64_t wseed, int wmin, int wmax, int threads, indigo_host_graph* out)
{
  if (out == nullptr || nodes < 1 || npairs < 0 || (npairs > 0 && (src == nullptr || dst == nullptr)))
    return set_error(INDIGO_ERR_INVALID_ARG, "build_csr: bad arguments");
  clear(out);
  const int nt = pick_threads(threads);
  for (long long i = 0; i < npairs; i++) {
    if (src[i] < 0 || src[i] >= nodes || dst[i] < 0 || dst[i] >= nodes)
      return set_error(INDIGO_ERR_INVALID_ARG, "build_csr: vertex id out of range");
  }

  // Bucket the pairs by source.  With room for one count array per thread,
  // every thread counts and later places its own share of the pairs through
  // private cursors (no atomics: hub vertices would make them contend);
  // otherwise shared atomic counters.  Lists are sorted afterwards, so the
  // order of placement does not show in the result.
  std::vector<long long> off(static_cast<size_t>(nodes) + 1, 0);
  std::vector<int> buf;
  const int parts = static_cast<int>(std::min<long long>(nt, std::max<long long>(1, npairs / 4096 + 1)));
  if (parts > 1 && static_cast<long long>(parts) * nodes <= (1LL << 27)) {
    std::vector<std::vector<int>> cnt(static_cast<size_t>(parts));
    run_threads(parts, [&](int p) {
      std::vector<int>& c = cnt[static_cast<size_t>(p)];
      c.assign(static_cast<size_t>(nodes), 0);
      for (long long i = npairs * p / parts; i < npairs * (p + 1) / parts; i++) {
        const int a = src[i], b = dst[i];
        if (a == b) continue;
        c[a]++;
        if (symmetrize) c[b]++;
      }
    });
    std::vector<long long> tot(static_cast<size_t>(nodes));
    parallel_for(nodes, nt, [&](long long b, long long e, int) {
      for (long long v = b; v < e; v++) {
        long long t = 0;
        for (int p = 0; p < parts; p++) t += cnt[static_cast<size_t>(p)][v];
        tot[v] = t;
      }
    });
    for (int v = 0; v < nodes; v++) off[v + 1] = off[v] + tot[v];
    buf.resize(static_cast<size_t>(std::max<long long>(off[nodes], 1)));
    // per-part cursors: part p writes after the entries of parts 0..p-1; the
    // cursors (64-bit list offsets of up to 2^31 entries) replace the counts
    std::vector<std::vector<long long>> cur(static_cast<size_t>(parts));
    run_threads(parts, [&](int p) { cur[static_cast<size_t>(p)].resize(static_cast<size_t>(nodes)); });
    parallel_for(nodes, nt, [&](long long b, long long e, int) {
      for (long long v = b; v < e; v++) {
        long long at = off[v];
        for (int p = 0; p < parts; p++) {
          cur[static_cast<size_t>(p)][v] = at;
          at += cnt[static_cast<size_t>(p)][v];
        }
      }
    });
    cnt.clear();
    run_threads(parts, [&](int p) {
      std::vector<long long>& c = cur[static_cast<size_t>(p)];
      for (long long i = npairs * p / parts; i < npairs * (p + 1) / parts; i++) {
        const int a = src[i], b = dst[i];
        if (a == b) continue;
        buf[static_cast<size_t>(c[a]++)] = b;
        if (symmetrize) buf[static_cast<size_t>(c[b]++)] = a;
      }
    });
  } else {
    std::vector<std::atomic<long long>> cnt(static_cast<size_t>(nodes) + 1);
    for (auto& c : cnt) c.store(0, std::memory_order_relaxed);
    parallel_for(npairs, nt, [&](long long b, long long e, int) {
      for (long long i = b; i < e; i++) {
        const int a = src[i], d = dst[i];
        if (a == d) continue;
        cnt[a].fetch_add(1, std::memory_order_relaxed);
        if (symmetrize) cnt[d].fetch_add(1, std::memory_order_relaxed);
      }
    });
    for (int v = 0; v < nodes; v++) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
    buf.resize(static_cast<size_t>(std::max<long long>(off[nodes], 1)));
    for (int v = 0; v < nodes; v++) cnt[v].store(off[v], std::memory_order_relaxed);
    parallel_for(npairs, nt, [&](long long b, long long e, int) {
      for (long long i = b; i < e; i++) {
        const int a = src[i], d = dst[i];
        if (a == d) continue;
        buf[cnt[a].fetch_add(1, std::memory_order_relaxed)] = d;
        if (symmetrize) buf[cnt[d].fetch_add(1, std::memory_order_relaxed)] = a;
      }
    });
  }
  // sort + unique each list; new degree into deg[].  Vertices are handed out
  // in chunks from a shared counter: R-MAT puts its long lists at small ids,
  // so static ranges would leave one thread with all of them.
  std::vector<int> deg(nodes);
  {
    constexpr int kChunk = 2048;
    std::atomic<long long> next{0};
    run_threads(std::min<int>(nt, nodes / kChunk + 1), [&](int) {
      for (;;) {
        const long long b = next.fetch_add(kChunk, std::memory_order_relaxed);
        if (b >= nodes) break;
        const long long e = std::min<long long>(nodes, b + kChunk);
        for (long long v = b; v < e; v++) {
          int* first = buf.data() + off[v];
          int* last = buf.data() + off[v + 1];
          std::sort(first, last);
          deg[v] = static_cast<int>(std::unique(first, last) - first);
        }
      }
    });
  }
  long long total = 0;
  for (int v = 0; v < nodes; v++) total += deg[v];
  if (total > 0x7fffffffLL) return set_error(INDIGO_ERR_CAPACITY, "build_csr: more than 2^31-1 CSR entries");

  out->nodes = nodes;
  out->edges = static_cast<int>(total);
  out->nindex = static_cast<int*>(malloc(sizeof(int) * (static_cast<size_t>(nodes) + 1)));
  out->nlist = static_cast<int*>(malloc(sizeof(int) * static_cast<size_t>(std::max<long long>(total, 1))));
  const bool want_w = weights_wanted(wmin, wmax);
  out->eweight = want_w ? static_cast<int*>(malloc(sizeof(int) * static_cast<size_t>(std::max<long long>(total, 1)))) : nullptr;
  if (!out->nindex || !out->nlist || (want_w && !out->eweight)) {
    indigo_graph_release(out);
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "build_csr: host allocation failed");
  }
  out->nindex[0] = 0;
  for (int v = 0; v < nodes; v++) out->nindex[v + 1] = out->nindex[v] + deg[v];
  parallel_for(nodes, nt, [&](long long b, long long e, int) {
    for (long long v = b; v < e; v++) {
      const int o = out->nindex[v];
      std::memcpy(out->nlist + o, buf.data() + off[v], sizeof(int) * deg[v]);
      if (want_w)
        for (int j = 0; j < deg[v]; j++) out->eweight[o + j] = edge_weight(wseed, static_cast<int>(v), out->nlist[o + j], wmin, wmax);
    }
  });
  return INDIGO_OK;
}

}  // namespace

extern "C" {

int indigo_build_csr(int nodes, long long npairs, const int* src, const int* dst, int symmetrize,
                     unsigned long long wseed, int wmin, int wmax, int threads, indigo_host_graph* out)
{
  return build_csr_impl(nodes, npairs, src, dst, symmetrize != 0, wseed, wmin, wmax, threads, out);
}

int indigo_gen_grid(int rows, int cols, unsigned long long seed, int wmin, int wmax, indigo_host_graph* out)
{
  if (out == nullptr || rows < 1 || cols < 1 || static_cast<long long>(rows) * cols > 0x7fffffffLL / 5)
    return set_error(INDIGO_ERR_INVALID_ARG, "gen_grid: bad size");
  clear(out);
  const int n = rows * cols;
  long long e = 0;
  std::vector<int> deg(n);
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++)
      deg[r * cols + c] = (r > 0) + (c > 0) + (c + 1 < cols) + (r + 1 < rows);
  for (int v = 0; v < n; v++) e += deg[v];
  out->nodes = n;
  out->edges = static_cast<int>(e);
  out->nindex = static_cast<int*>(malloc(sizeof(int) * (static_cast<size_t>(n) + 1)));
  out->nlist = static_cast<int*>(malloc(sizeof(int) * static_cast<size_t>(std::max<long long>(e, 1))));
  const bool want_w = weights_wanted(wmin, wmax);
  out->eweight = want_w ? static_cast<int*>(malloc(sizeof(int) * static_cast<size_t>(std::max<long long>(e, 1)))) : nullptr;
  if (!out->nindex || !out->nlist || (want_w && !out->eweight)) {
    indigo_graph_release(out);
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "gen_grid: host allocation failed");
  }
  int k = 0;
  out->nindex[0] = 0;
  for (int r = 0; r < rows; r++) {
    for (int c = 0; c < cols; c++) {
      const int v = r * cols + c;
      // ascending ids: up, left, right, down
      if (r > 0) out->nlist[k++] = v - cols;
      if (c > 0) out->nlist[k++] = v - 1;
      if (c + 1 < cols) out->nlist[k++] = v + 1;
      if (r + 1 < rows) out->nlist[k++] = v + cols;
      out->nindex[v + 1] = k;
    }
  }
  if (want_w)
    for (int v = 0; v < n; v++)
      for (int j = out->nindex[v]; j < out->nindex[v + 1]; j++) out->eweight[j] = edge_weight(seed, v, out->nlist[j], wmin, wmax);
  return INDIGO_OK;
}

int indigo_gen_rmat(int scale, int edge_factor, double a, double b, double c, unsigned long long seed,
                    int wmin, int wmax, int threads, indigo_host_graph* out)
{
  if (out == nullptr || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return set_error(INDIGO_ERR_INVALID_ARG, "gen_rmat: bad parameters");
  const long long n = 1LL << scale;
  const long long samples = static_cast<long long>(edge_factor) * n;
  if (2 * samples > 0x7fffffffLL * 2) return set_error(INDIGO_ERR_CAPACITY, "gen_rmat: too many samples");
  std::vector<int> s(samples), d(samples);
  const int nt = pick_threads(threads);
  const double ab = a + b, abc = a + b + c;
  parallel_for(samples, nt, [&](long long beg, long long end, int) {
    for (long long i = beg; i < end; i++) {
      Stream rng(seed, static_cast<uint64_t>(i));
      int u = 0, v = 0;
      for (int lvl = 0; lvl < scale; lvl++) {
        const double r = rng.uniform();
        const int ubit = r >= ab;                       // c or d quadrant -> lower half rows
        const int vbit = (r >= a && r < ab) || r >= abc;  // b or d quadrant -> right half cols
        u = (u << 1) | ubit;
        v = (v << 1) | vbit;
      }
      s[i] = u;
      d[i] = v;
    }
  });
  return build_csr_impl(static_cast<int>(n), samples, s.data(), d.data(), true, seed ^ 0x5eedULL, wmin, wmax, threads, out);
}

int indigo_gen_er(int nodes, long long samples, unsigned long long seed, int wmin, int wmax, int threads,
                  indigo_host_graph* out)
{
  if (out == nullptr || nodes < 1 || samples < 0) return set_error(INDIGO_ERR_INVALID_ARG, "gen_er: bad parameters");
  std::vector<int> s(samples), d(samples);
  const int nt = pick_threads(threads);
  parallel_for(samples, nt, [&](long long beg, long long end, int) {
    for (long long i = beg; i < end; i++) {
      Stream rng(seed, static_cast<uint64_t>(i));
      s[i] = static_cast<int>(rng.next() % static_cast<uint64_t>(nodes));
      d[i] = static_cast<int>(rng.next() % static_cast<uint64_t>(nodes));
    }
  });
  return build_csr_impl(nodes, samples, s.data(), d.data(), true, seed ^ 0x5eedULL, wmin, wmax, threads, out);
}

int indigo_gen_road(int rows, int cols, double keep_p, double diag_p, unsigned long long seed, int wmin,
                    int wmax, int threads, indigo_host_graph* out)
{
  if (out == nullptr || rows < 1 || cols < 1 || keep_p < 0 || keep_p > 1 || diag_p < 0 || diag_p > 1 ||
      static_cast<long long>(rows) * cols > 0x7fffffffLL / 8)
    return set_error(INDIGO_ERR_INVALID_ARG, "gen_road: bad parameters");
  const long long n = static_cast<long long>(rows) * cols;
  // per vertex: right edge, down edge, optional diagonal -> at most 3 pairs
  std::vector<int> s(3 * n), d(3 * n);
  std::vector<unsigned char> used(3 * n, 0);
  const int nt = pick_threads(threads);
  parallel_for(n, nt, [&](long long beg, long long end, int) {
    for (long long v = beg; v < end; v++) {
      Stream rng(seed, static_cast<uint64_t>(v));
      const int r = static_cast<int>(v / cols), c = static_cast<int>(v % cols);
      const double pr = rng.uniform(), pd = rng.uniform(), pg = rng.uniform();
      const uint64_t pick = rng.next();
      if (c + 1 < cols && pr < keep_p) { s[3 * v] = static_cast<int>(v); d[3 * v] = static_cast<int>(v + 1); used[3 * v] = 1; }
      if (r + 1 < rows && pd < keep_p) { s[3 * v + 1] = static_cast<int>(v); d[3 * v + 1] = static_cast<int>(v + cols); used[3 * v + 1] = 1; }
      if (pg < diag_p) {
        int cand[4], m = 0;
        const int dr[4] = {-1, -1, 1, 1}, dc[4] = {-1, 1, -1, 1};
        for (int k = 0; k < 4; k++) {
          const int rr = r + dr[k], cc = c + dc[k];
          if (rr >= 0 && rr < rows && cc >= 0 && cc < cols) cand[m++] = rr * cols + cc;
        }
        if (m > 0) { s[3 * v + 2] = static_cast<int>(v); d[3 * v + 2] = cand[pick % m]; used[3 * v + 2] = 1; }
      }
    }
  });
  long long k = 0;
  for (long long i = 0; i < 3 * n; i++) {
    if (used[i]) { s[k] = s[i]; d[k] = d[i]; k++; }
  }
  return build_csr_impl(static_cast<int>(n), k, s.data(), d.data(), true, seed ^ 0x5eedULL, wmin, wmax, threads, out);
}

int indigo_graph_copy(const indigo_host_graph* src, indigo_host_graph* out)
{
  if (src == nullptr || out == nullptr || src->nodes < 1 || src->edges < 0)
    return set_error(INDIGO_ERR_INVALID_ARG, "graph_copy: bad arguments");
  clear(out);
  out->nodes = src->nodes;
  out->edges = src->edges;
  const size_t ne = static_cast<size_t>(std::max(src->edges, 1));
  out->nindex = static_cast<int*>(malloc(sizeof(int) * (static_cast<size_t>(src->nodes) + 1)));
  out->nlist = static_cast<int*>(malloc(sizeof(int) * ne));
  out->eweight = src->eweight ? static_cast<int*>(malloc(sizeof(int) * ne)) : nullptr;
  if (!out->nindex || !out->nlist || (src->eweight && !out->eweight)) {
    indigo_graph_release(out);
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "graph_copy: host allocation failed");
  }
  std::memcpy(out->nindex, src->nindex, sizeof(int) * (static_cast<size_t>(src->nodes) + 1));
  std::memcpy(out->nlist, src->nlist, sizeof(int) * static_cast<size_t>(src->edges));
  if (src->eweight) std::memcpy(out->eweight, src->eweight, sizeof(int) * static_cast<size_t>(src->edges));
  return INDIGO_OK;
}

int indigo_graph_reorder(const indigo_host_graph* src, int threads, indigo_host_graph* out, int* perm)
{
  if (src == nullptr || out == nullptr || perm == nullptr || src->nodes < 1 || src->edges < 0 ||
      src->nindex == nullptr || (src->edges > 0 && src->nlist == nullptr))
    return set_error(INDIGO_ERR_INVALID_ARG, "graph_reorder: bad arguments");
  const int n = src->nodes;
  const int nt = pick_threads(threads);
  clear(out);
  // new ids by degree, largest first, ties in id order: a counting sort
  // over the degrees, stable because vertices are placed in id order
  int maxdeg = 0;
  for (int v = 0; v < n; v++) {
    const int d = src->nindex[v + 1] - src->nindex[v];
    if (d < 0 || src->nindex[v] < 0 || src->nindex[v + 1] > src->edges)
      return set_error(INDIGO_ERR_INVALID_ARG, "graph_reorder: nindex is not a CSR offset array");
    maxdeg = std::max(maxdeg, d);
  }
  std::vector<long long> start(static_cast<size_t>(maxdeg) + 2, 0);
  for (int v = 0; v < n; v++) start[static_cast<size_t>(maxdeg - (src->nindex[v + 1] - src->nindex[v])) + 1]++;
  for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
  std::vector<int> order(static_cast<size_t>(n));
  for (int v = 0; v < n; v++) {
    const size_t k = static_cast<size_t>(maxdeg - (src->nindex[v + 1] - src->nindex[v]));
    const int id = static_cast<int>(start[k]++);
    perm[v] = id;
    order[static_cast<size_t>(id)] = v;
  }
  out->nodes = n;
  out->edges = src->edges;
  const size_t ne = static_cast<size_t>(std::max(src->edges, 1));
  out->nindex = static_cast<int*>(malloc(sizeof(int) * (static_cast<size_t>(n) + 1)));
  out->nlist = static_cast<int*>(malloc(sizeof(int) * ne));
  out->eweight = src->eweight ? static_cast<int*>(malloc(sizeof(int) * ne)) : nullptr;
  if (!out->nindex || !out->nlist || (src->eweight && !out->eweight)) {
    indigo_graph_release(out);
    return set_error(INDIGO_ERR_OUT_OF_MEMORY, "graph_reorder: host allocation failed");
  }
  out->nindex[0] = 0;
  for (int i = 0; i < n; i++) {
    const int v = order[static_cast<size_t>(i)];
    out->nindex[i + 1] = out->nindex[i] + (src->nindex[v + 1] - src->nindex[v]);
  }
  // rows: each vertex's list renamed and sorted by new id (weights moved
  // with their entries); rows handed out in chunks, long ones come first
  std::atomic<bool> bad{false};
  {
    constexpr int kChunk = 1024;
    std::atomic<long long> next{0};
    run_threads(std::min<int>(nt, n / kChunk + 1), [&](int) {
      std::vector<std::pair<int, int>> row;
      for (;;) {
        const long long b = next.fetch_add(kChunk, std::memory_order_relaxed);
        if (b >= n) break;
        const long long e = std::min<long long>(n, b + kChunk);
        for (long long i = b; i < e; i++) {
          const int v = order[static_cast<size_t>(i)];
          const int o = src->nindex[v], d = src->nindex[v + 1] - o;
          row.resize(static_cast<size_t>(d));
          for (int j = 0; j < d; j++) {
            const int u = src->nlist[o + j];
            if (u < 0 || u >= n) {
              bad.store(true, std::memory_order_relaxed);
              row[static_cast<size_t>(j)] = {0, 0};
              continue;
            }
            row[static_cast<size_t>(j)] = {perm[u], src->eweight ? src->eweight[o + j] : 0};
          }
          std::sort(row.begin(), row.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
            return x.first < y.first;
          });
          const int q = out->nindex[i];
          for (int j = 0; j < d; j++) {
            out->nlist[q + j] = row[static_cast<size_t>(j)].first;
            if (out->eweight) out->eweight[q + j] = row[static_cast<size_t>(j)].second;
          }
        }
      }
    });
  }
  if (bad.load()) {
    indigo_graph_release(out);
    return set_error(INDIGO_ERR_INVALID_ARG, "graph_reorder: a list entry is not a vertex id");
  }
  return INDIGO_OK;
}

int indigo_graph_release(indigo_host_graph* g)
{
  if (g == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "graph_release: null");
  free(g->nindex);
  free(g->nlist);
  free(g->eweight);
  clear(g);
  return INDIGO_OK;
}

int indigo_read_egr(const char* path, indigo_host_graph* out)
{
  if (path == nullptr || out == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "read_egr: null argument");
  clear(out);
  FILE* f = fopen(path, "rb");
  if (f == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "read_egr: could not open file");
  int hdr[2];
  if (fread(hdr, sizeof(int), 2, f) != 2 || hdr[0] < 1 || hdr[1] < 0) {
    fclose(f);
    return set_error(INDIGO_ERR_INVALID_ARG, "read_egr: bad header");
  }
  out->nodes = hdr[0];
  out->edges = hdr[1];
  const size_t nv = static_cast<size_t>(hdr[0]) + 1, ne = static_cast<size_t>(hdr[1]);
  out->nindex = static_cast<int*>(malloc(sizeof(int) * nv));
  out->nlist = static_cast<int*>(malloc(sizeof(int) * std::max<size_t>(ne, 1)));
  out->eweight = static_cast<int*>(malloc(sizeof(int) * std::max<size_t>(ne, 1)));
  int rc = INDIGO_OK;
  if (!out->nindex || !out->nlist || !out->eweight) rc = set_error(INDIGO_ERR_OUT_OF_MEMORY, "read_egr: allocation failed");
  else if (fread(out->nindex, sizeof(int), nv, f) != nv) rc = set_error(INDIGO_ERR_INVALID_ARG, "read_egr: truncated nindex");
  else if (fread(out->nlist, sizeof(int), ne, f) != ne) rc = set_error(INDIGO_ERR_INVALID_ARG, "read_egr: truncated nlist");
  else {
    const size_t got = ne ? fread(out->eweight, sizeof(int), ne, f) : 0;
    if (got == 0) {
      free(out->eweight);
      out->eweight = nullptr;
    } else if (got != ne) {
      rc = set_error(INDIGO_ERR_INVALID_ARG, "read_egr: truncated weights");
    }
  }
  fclose(f);
  if (rc != INDIGO_OK) indigo_graph_release(out);
  return rc;
}

int indigo_egr_open(const char* path, indigo_egr_map* out)
{
  if (path == nullptr || out == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "egr_open: null argument");
  std::memset(out, 0, sizeof(*out));
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return set_error(INDIGO_ERR_INVALID_ARG, "egr_open: could not open file");
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < static_cast<off_t>(2 * sizeof(int))) {
    close(fd);
    return set_error(INDIGO_ERR_INVALID_ARG, "egr_open: bad header");
  }
  const size_t len = static_cast<size_t>(st.st_size);
  void* base = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (base == MAP_FAILED) return set_error(INDIGO_ERR_OUT_OF_MEMORY, "egr_open: mmap failed");
  const int* w = static_cast<const int*>(base);
  const long long nodes = w[0], edges = w[1];
  // the reader's rules (include/ECLgraph.h:58-76): header, nindex, nlist;
  // then no whole weight means no weights, some but fewer than `edges` is
  // an error, and bytes after the weights are ignored
  const size_t need = sizeof(int) * static_cast<size_t>(2 + (nodes + 1) + (edges > 0 ? edges : 0));
  const size_t extra_ints = len >= need ? (len - need) / sizeof(int) : 0;
  int rc = INDIGO_OK;
  if (nodes < 1 || edges < 0) rc = set_error(INDIGO_ERR_INVALID_ARG, "egr_open: bad header");
  else if (len < need) rc = set_error(INDIGO_ERR_INVALID_ARG, "egr_open: truncated nindex / nlist");
  else if (extra_ints > 0 && extra_ints < static_cast<size_t>(edges))
    rc = set_error(INDIGO_ERR_INVALID_ARG, "egr_open: truncated weights");
  if (rc != INDIGO_OK) {
    munmap(base, len);
    return rc;
  }
  madvise(base, len, MADV_SEQUENTIAL);
  out->nodes = static_cast<int>(nodes);
  out->edges = static_cast<int>(edges);
  out->nindex = w + 2;
  out->nlist = w + 2 + nodes + 1;
  out->eweight = (extra_ints > 0 && edges > 0) ? w + 2 + nodes + 1 + edges : nullptr;
  out->base = base;
  out->length = len;
  return INDIGO_OK;
}

int indigo_egr_close(indigo_egr_map* m)
{
  if (m == nullptr) return INDIGO_OK;
  if (m->base != nullptr) munmap(m->base, m->length);
  std::memset(m, 0, sizeof(*m));
  return INDIGO_OK;
}

int indigo_write_egr(const char* path, const indigo_host_graph* g)
{
  if (path == nullptr || g == nullptr || g->nodes < 1 || g->edges < 0 || !g->nindex || (!g->nlist && g->edges))
    return set_error(INDIGO_ERR_INVALID_ARG, "write_egr: bad arguments");
  FILE* f = fopen(path, "wb");
  if (f == nullptr) return set_error(INDIGO_ERR_INVALID_ARG, "write_egr: could not open file");
  const int hdr[2] = {g->nodes, g->edges};
  bool ok = fwrite(hdr, sizeof(int), 2, f) == 2;
  ok = ok && fwrite(g->nindex, sizeof(int), static_cast<size_t>(g->nodes) + 1, f) == static_cast<size_t>(g->nodes) + 1;
  ok = ok && fwrite(g->nlist, sizeof(int), static_cast<size_t>(g->edges), f) == static_cast<size_t>(g->edges);
  if (g->eweight) ok = ok && fwrite(g->eweight, sizeof(int), static_cast<size_t>(g->edges), f) == static_cast<size_t>(g->edges);
  ok = (fclose(f) == 0) && ok;
  return ok ? INDIGO_OK : set_error(INDIGO_ERR_INVALID_ARG, "write_egr: write failed");
}

int indigo_graph_check(const indigo_host_graph* g)
{
  if (g == nullptr || g->nodes < 1 || g->edges < 0 || !g->nindex || (!g->nlist && g->edges > 0)) return 1;
  if (g->nindex[0] != 0 || g->nindex[g->nodes] != g->edges) return 1;
  for (int v = 0; v < g->nodes; v++)
    if (g->nindex[v + 1] < g->nindex[v]) return 1;
  for (int v = 0; v < g->nodes; v++) {
    for (int j = g->nindex[v]; j < g->nindex[v + 1]; j++) {
      const int u = g->nlist[j];
      if (u < 0 || u >= g->nodes) return 2;
      if (j > g->nindex[v] && g->nlist[j - 1] >= u) return 3;
      if (u == v) return 4;
      // reverse edge by binary search in u's (sorted) list
      const int* b = g->nlist + g->nindex[u];
      const int* e = g->nlist + g->nindex[u + 1];
      const int* p = std::lower_bound(b, e, v);
      if (p == e || *p != v) return 5;
      if (g->eweight && g->eweight[p - g->nlist] != g->eweight[j]) return 6;
    }
  }
  return 0;
}

}  // extern "C"
