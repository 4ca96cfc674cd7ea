// Serving loop over resident graphs: continuous batching on several slots.
//
// Requests (algorithm, graph, source, arrival time) are replayed in real
// time.  A slot is a context (own stream, own workspace) with a 1/slots share
// of the persistent grid.  Whenever a slot is free it takes, in arrival
// order, the waiting requests whose outputs fit the memory budget (device
// arena, first fit): BFS and SSSP requests on the same graph become one
// multi-source task (up to 64 / 32 sources), CC, MIS and PR one task each,
// all in one batched launch (indigo_batch); a triangle count at the head of
// the queue takes the slot alone (its kernels are ordinary launches), and so
// does a BFS or SSSP request on a deep, sparse graph (average degree below
// 6, at least 65536 vertices: grids, road networks), on the single-query
// kernels: their one-workgroup phase runs the thousands of narrow levels at
// workgroup-barrier cost (the batched engine's gangs pay a device barrier
// per level: 82 ms against about 45 for a 2048^2 grid's BFS), and the
// batched Bellman-Ford rounds re-relax such graphs many times over, while
// near-far settles them in distance order.  The
// other slots keep running, so requests that arrive while the GPU is busy
// are admitted as soon as a slot and memory are free, and their work
// overlaps on the device.  Completion is polled with events; results are
// copied to the caller's buffers and the arena space released.
//
// Head-of-line blocking.  A launch returns all its results together, so a
// short request batched with a long one waits for the long one, and short
// requests queued behind long launches on every slot wait for all of them.
// Each request therefore gets an estimated cost (est_us: its graph's CSR
// entries at the algorithm's measured rate, or its level count on deep
// sparse graphs) and a class: short (at most kLongUs) or long.  A launch
// holds requests of one class only; long work never occupies the last free
// slot (one slot stays reserved for short work when there are two or more);
// and a free slot serves the class of the oldest waiting request, or short
// work when long work may not take it - short requests overtake long ones,
// never the other way round.  Slots are capped at the process's hardware
// queues (GPU_MAX_HW_QUEUES, 4 when unset): more streams than queues share
// queues and serialise instead of overlapping.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "server.hpp"

#include "batch.hpp"
#include "bfs.hpp"
#include "sssp.hpp"
#include "runtime.hpp"
#include "tc.hpp"

namespace indigo {

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;

size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over one device buffer.
class Arena {
 public:
  ~Arena()
  {
    if (base_ != nullptr) (void)hipFree(base_);
  }
  int init(size_t bytes)
  {
    size_ = round_up(bytes);
    IG_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&base_), size_));
    free_[0] = size_;
    return INDIGO_OK;
  }
  bool alloc(size_t bytes, size_t* off)
  {
    bytes = round_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t o = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest > 0) free_[o + bytes] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      *off = o;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes)
  {
    bytes = round_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t free_bytes() const { return size_ - used_; }
  size_t size() const { return size_; }
  size_t peak() const { return peak_; }
  char* ptr(size_t off) const { return base_ + off; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

struct Piece {
  size_t off, bytes;
};

struct Route {
  int req;
  const void* src;  // result in the arena
  size_t bytes;
};

// Requests estimated to run longer than this (microseconds on the whole
// GPU) are long.
constexpr double kLongUs = 8000.0;  // estimated slot microseconds above which a request is long
constexpr double kBatchWorkUs = 1e30;  // default: no cap on a launch's estimated work

struct Slot {
  indigo_ctx* ctx = nullptr;
  int state = 0;  // 0 idle, 1 batch, 2 triangle count, 3 single BFS or SSSP
  bool long_work = false;  // the running launch holds long requests
  std::vector<int> reqs;
  std::vector<Piece> pieces;
  std::vector<indigo_query> queries;
  std::vector<Route> routes;
  int tc_req = -1;
  int single_req = -1;
  BfsStats bfs_stats;
};

struct Server {
  const indigo_graph_ref* graphs;
  int ngraphs;
  indigo_request* reqs;
  int nreqs;
  indigo_serve_config cfg;
  indigo_serve_stats stats{};
  Arena arena;
  std::vector<Slot> slots;
  std::vector<int> waiting;  // arrived, not admitted (arrival order)
  int grid_per_slot = 1;
  Clock::time_point t0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  bool single(const indigo_request& r) const
  {
    const indigo_graph_ref& G = graphs[r.graph];
    return (r.alg == INDIGO_QUERY_SSSP || r.alg == INDIGO_QUERY_BFS) && G.nodes >= (1 << 16) &&
           static_cast<long long>(G.edges) < 6LL * G.nodes;
  }

  bool deep(const indigo_request& r) const
  {
    const indigo_graph_ref& G = graphs[r.graph];
    return G.nodes >= (1 << 16) && static_cast<long long>(G.edges) < 6LL * G.nodes;
  }

  // Estimated microseconds of one request's service on a slot (its share of
  // the GPU, beside the other slots), fitted to the median service times of
  // the 64-graph serving pool at 150 requests/s: a fixed part (launch, grid
  // assembly, gang barriers; PageRank's sweeps make it 13 ms) plus CSR
  // entries at a per-algorithm rate (entries per microsecond: BFS 10000,
  // SSSP 3000, CC 50000, MIS 10000, PageRank 1600, TC 2000).  BFS and SSSP
  // on deep sparse graphs run alone on the single-query kernels and are
  // costed by their level count instead (about 2 sqrt(V) levels from a
  // random source at 10 and 20 us per level).
  double est_us(const indigo_request& r) const
  {
    const indigo_graph_ref& G = graphs[r.graph];
    const double e = static_cast<double>(G.edges);
    if (deep(r) && (r.alg == INDIGO_QUERY_BFS || r.alg == INDIGO_QUERY_SSSP)) {
      const double levels = 2.0 * std::sqrt(static_cast<double>(G.nodes));
      return levels * (r.alg == INDIGO_QUERY_BFS ? 10.0 : 20.0);
    }
    double fixed = 2000.0, rate = 10000.0;
    switch (r.alg) {
      case INDIGO_QUERY_SSSP: rate = 3000.0; break;
      case INDIGO_QUERY_CC: rate = 50000.0; break;
      case INDIGO_QUERY_PR: fixed = 13000.0; rate = 1600.0; break;
      case INDIGO_QUERY_TC: rate = 2000.0; break;
      default: break;
    }
    return fixed + e / rate;
  }

  bool is_long(const indigo_request& r) const { return est_us(r) > kLongUs; }

  size_t out_bytes(const indigo_request& r) const
  {
    const size_t v = static_cast<size_t>(graphs[r.graph].nodes);
    return r.alg == INDIGO_QUERY_PR ? 4 * (v + 1) : 4 * v;
  }

  // Takes waiting requests of one class (long or short) into one batched
  // launch on slot s; false if none fit.
  int dispatch_batch(Slot& s, bool want_long, bool* launched)
  {
    *launched = false;
    struct Build {
      int alg, graph;
      std::vector<int> reqs;
    };
    std::vector<Build> builds;
    std::map<std::pair<int, int>, int> open;  // (alg, graph) -> build with free source slots
    std::vector<char> taken(waiting.size(), 0);
    size_t need = 0;
    int count = 0;
    double work = 0.0;  // estimated whole-GPU microseconds taken so far
    const double work_cap = cfg.batch_work_us > 0 ? cfg.batch_work_us : kBatchWorkUs;
    for (size_t i = 0; i < waiting.size() && count < cfg.max_batch; i++) {
      const indigo_request& r = reqs[waiting[i]];
      if (r.alg == INDIGO_QUERY_TC || single(r) || is_long(r) != want_long) continue;
      const size_t b = round_up(out_bytes(r));
      if (need + b > arena.free_bytes()) break;  // within a class, arrival order: no overtaking
      const bool multi = r.alg == INDIGO_QUERY_BFS || r.alg == INDIGO_QUERY_SSSP;
      const int cap = r.alg == INDIGO_QUERY_BFS ? INDIGO_MAX_BFS_SOURCES : INDIGO_MAX_SSSP_SOURCES;
      auto it = multi ? open.find({r.alg, r.graph}) : open.end();
      const bool joins = it != open.end() && static_cast<int>(builds[it->second].reqs.size()) < cap;
      // one more source of a multi-source task shares its traversal: a
      // quarter of a request's estimate
      const double w = joins ? 0.25 * est_us(r) : est_us(r);
      if (count > 0 && work + w > work_cap) break;  // within a class, arrival order: no overtaking
      work += w;
      if (joins) {
        builds[it->second].reqs.push_back(waiting[i]);
      } else {
        builds.push_back(Build{r.alg, r.graph, {waiting[i]}});
        if (multi) open[{r.alg, r.graph}] = static_cast<int>(builds.size()) - 1;
      }
      need += b;
      taken[i] = 1;
      count++;
    }
    if (count == 0) return INDIGO_OK;
    // one arena block per task (multi-source slabs are contiguous)
    s.queries.clear();
    s.routes.clear();
    s.pieces.clear();
    s.reqs.clear();
    for (const Build& bd : builds) {
      const indigo_request& r0 = reqs[bd.reqs[0]];
      const size_t per = out_bytes(r0);
      size_t off = 0;
      if (!arena.alloc(per * bd.reqs.size(), &off)) {
        // fragmentation: give this task's requests back (they stay waiting)
        for (int q : bd.reqs)
          for (size_t i = 0; i < waiting.size(); i++)
            if (waiting[i] == q) taken[i] = 0;
        continue;
      }
      s.pieces.push_back(Piece{off, per * bd.reqs.size()});
      indigo_query q{};
      q.alg = bd.alg;
      q.graph = bd.graph;
      q.nsrc = static_cast<int>(bd.reqs.size());
      for (size_t k = 0; k < bd.reqs.size(); k++) q.src[k] = reqs[bd.reqs[k]].src;
      q.d_out = reinterpret_cast<int*>(arena.ptr(off));
      s.queries.push_back(q);
      for (size_t k = 0; k < bd.reqs.size(); k++) {
        s.reqs.push_back(bd.reqs[k]);
        s.routes.push_back(Route{bd.reqs[k], arena.ptr(off) + k * per, per});
      }
    }
    if (s.queries.empty()) return INDIGO_OK;
    std::vector<int> rest;
    for (size_t i = 0; i < waiting.size(); i++)
      if (!taken[i]) rest.push_back(waiting[i]);
    waiting.swap(rest);
    const double t = now();
    for (int q : s.reqs) {
      reqs[q].admitted = t;
      reqs[q].batch = static_cast<int>(s.reqs.size());
      reqs[q].slot = static_cast<int>(&s - slots.data());
    }
    IG_TRY(batch_launch(s.ctx, graphs, ngraphs, s.queries.data(), static_cast<int>(s.queries.size()),
                        grid_per_slot));
    s.state = 1;
    s.long_work = want_long;
    stats.launches++;
    *launched = true;
    return INDIGO_OK;
  }

  int dispatch_tc(Slot& s, int q)
  {
    const indigo_graph_ref& G = graphs[reqs[q].graph];
    IG_TRY(ensure_workspace(s.ctx, tc_workspace_bytes(G.nodes, G.edges)));
    IG_TRY(ensure_host_pinned(s.ctx, 4096));
    waiting.erase(std::find(waiting.begin(), waiting.end(), q));
    reqs[q].admitted = now();
    reqs[q].batch = 1;
    reqs[q].slot = static_cast<int>(&s - slots.data());
    IG_TRY(timing_begin(s.ctx));
    int launches = 0;
    IG_TRY(tc_run(s.ctx, G.nodes, G.edges, G.d_nindex, G.d_nlist, &launches));
    IG_TRY(timing_record_stop(s.ctx));
    s.tc_req = q;
    s.state = 2;
    s.long_work = is_long(reqs[q]);
    stats.tc_launches++;
    return INDIGO_OK;
  }

  // A BFS or SSSP request on a deep graph alone on slot s; launched false if
  // its output does not fit yet.
  int dispatch_single(Slot& s, int q, bool* launched)
  {
    *launched = false;
    const indigo_request& r = reqs[q];
    size_t off = 0;
    if (!arena.alloc(out_bytes(r), &off)) return INDIGO_OK;
    waiting.erase(std::find(waiting.begin(), waiting.end(), q));
    reqs[q].admitted = now();
    reqs[q].batch = 1;
    reqs[q].slot = static_cast<int>(&s - slots.data());
    s.pieces.assign(1, Piece{off, out_bytes(r)});
    s.routes.assign(1, Route{q, arena.ptr(off), out_bytes(r)});
    s.reqs.assign(1, q);
    s.single_req = q;
    s.long_work = is_long(r);
    IG_TRY(ensure_workspace(s.ctx, single_workspace_bytes(r)));
    IG_TRY(ensure_host_pinned(s.ctx, 4096));
    IG_TRY(timing_begin(s.ctx));
    IG_TRY(run_single(s, r, reinterpret_cast<int*>(arena.ptr(off))));
    IG_TRY(timing_record_stop(s.ctx));
    s.state = 3;
    stats.launches++;
    *launched = true;
    return INDIGO_OK;
  }

  size_t single_workspace_bytes(const indigo_request& r) const
  {
    const int n = graphs[r.graph].nodes;
    return r.alg == INDIGO_QUERY_BFS ? bfs_workspace_bytes(n) : sssp_workspace_bytes(n);
  }

  int run_single(Slot& s, const indigo_request& r, int* out)
  {
    const indigo_graph_ref& G = graphs[r.graph];
    if (r.alg == INDIGO_QUERY_BFS) {
      s.bfs_stats = BfsStats{};
      return bfs_run(s.ctx, G.nodes, G.edges, G.d_nindex, G.d_nlist, r.src, out, &s.bfs_stats);
    }
    return sssp_run(s.ctx, G.nodes, G.edges, G.d_nindex, G.d_nlist, G.d_eweight, r.src, out);
  }

  int collect_single(Slot& s, const indigo_request& r)
  {
    if (r.alg == INDIGO_QUERY_BFS) return bfs_collect(s.ctx, &s.bfs_stats);
    int it = 0, ep = 0;
    return sssp_collect(s.ctx, &it, &ep);
  }

  int complete(Slot& s, int* done)
  {
    if (s.state == 3) {
      const indigo_request& r = reqs[s.single_req];
      const int nl = r.alg == INDIGO_QUERY_BFS ? s.bfs_stats.launches : 2;
      IG_TRY(timing_finish(s.ctx, nl, 0));
      int rc = collect_single(s, r);
      while (rc == kNotResident) {  // the grid did not assemble: again, with the lowered cap
        IG_TRY(timing_begin(s.ctx));
        IG_TRY(run_single(s, r, reinterpret_cast<int*>(const_cast<void*>(s.routes[0].src))));
        IG_TRY(timing_end(s.ctx, r.alg == INDIGO_QUERY_BFS ? s.bfs_stats.launches : 2, 0));
        rc = collect_single(s, r);
      }
      IG_TRY(rc);
      return finish_routes(s, done);
    }
    if (s.state == 1) {
      IG_TRY(batch_wait(s.ctx));
      return finish_routes(s, done);
    }
    IG_TRY(timing_finish(s.ctx, 0, 1));
    long long count = 0;
    IG_TRY(tc_collect(s.ctx, &count));
    indigo_request& r = reqs[s.tc_req];
    if (r.d_out != nullptr) IG_HIP_TRY(hipMemcpy(r.d_out, &count, sizeof(count), hipMemcpyHostToDevice));
    r.finished = now();
    *done += 1;
    s.state = 0;
    return INDIGO_OK;
  }

  int finish_routes(Slot& s, int* done)
  {
    {
      bool copies = false;
      for (const Route& rt : s.routes) {
        if (reqs[rt.req].d_out == nullptr) continue;
        IG_HIP_TRY(hipMemcpyAsync(reqs[rt.req].d_out, rt.src, rt.bytes, hipMemcpyDeviceToDevice, s.ctx->stream));
        copies = true;
      }
      if (copies) IG_HIP_TRY(hipStreamSynchronize(s.ctx->stream));
      const double t = now();
      for (int q : s.reqs) reqs[q].finished = t;
      for (const Piece& p : s.pieces) arena.release(p.off, p.bytes);
      *done += static_cast<int>(s.reqs.size());
    }
    s.state = 0;
    return INDIGO_OK;
  }

  int run()
  {
    std::vector<int> order(static_cast<size_t>(nreqs));
    for (int i = 0; i < nreqs; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
    int next = 0, done = 0;
    t0 = Clock::now();
    double next_report = cfg.progress_interval;
    while (done < nreqs) {
      const double t = now();
      if (cfg.progress_interval > 0.0 && t >= next_report) {
        fprintf(stderr, "serve: %d of %d requests done, %zu waiting, %.1f s\n", done, nreqs, waiting.size(), t);
        fflush(stderr);
        next_report = t + cfg.progress_interval;
      }
      while (next < nreqs && reqs[order[next]].arrival <= t) waiting.push_back(order[next++]);
      stats.max_waiting = std::max(stats.max_waiting, static_cast<int>(waiting.size()));
      bool progressed = false;
      for (Slot& s : slots) {
        if (s.state == 0) continue;
        const hipError_t e = hipEventQuery(s.ctx->ev_stop);
        if (e == hipErrorNotReady) continue;
        IG_HIP_TRY(e);
        IG_TRY(complete(s, &done));
        progressed = true;
      }
      for (Slot& s : slots) {
        if (s.state != 0 || waiting.empty()) continue;
        // the request to serve: the oldest one, unless it is long and long
        // work may not take this slot - then the oldest short one
        int busy_long = 0;
        for (const Slot& o : slots) busy_long += o.state != 0 && o.long_work;
        const bool long_ok = slots.size() < 2 || busy_long < static_cast<int>(slots.size()) - 1;
        int pick = -1;
        for (size_t i = 0; i < waiting.size() && pick < 0; i++)
          if (long_ok || !is_long(reqs[waiting[i]])) pick = waiting[i];
        if (pick < 0) break;  // only long work waits, and it may not take another slot
        const indigo_request& head = reqs[pick];
        if (head.alg == INDIGO_QUERY_TC) {
          IG_TRY(dispatch_tc(s, pick));
          progressed = true;
          continue;
        }
        if (single(head)) {
          bool launched = false;
          IG_TRY(dispatch_single(s, pick, &launched));
          if (!launched) break;  // its output does not fit until something completes
          progressed = true;
          continue;
        }
        bool launched = false;
        IG_TRY(dispatch_batch(s, is_long(head), &launched));
        if (!launched) break;  // nothing fits until something completes
        progressed = true;
      }
      if (!progressed) {
        bool busy = false;
        for (const Slot& s : slots) busy |= s.state != 0;
        if (!busy && !waiting.empty()) {
          const indigo_request& r = reqs[waiting.front()];
          if (r.alg != INDIGO_QUERY_TC && round_up(out_bytes(r)) > arena.size())
            return set_error(INDIGO_ERR_CAPACITY, "serve: request output of %zu bytes exceeds the memory budget",
                             out_bytes(r));
        }
        std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
    stats.makespan = now();
    stats.peak_bytes = arena.peak();
    return INDIGO_OK;
  }
};

}  // namespace

int serve_run(const indigo_graph_ref* graphs, int ngraphs, indigo_request* reqs, int nreqs,
              const indigo_serve_config* cfg, indigo_serve_stats* stats)
{
  Server sv;
  sv.graphs = graphs;
  sv.ngraphs = ngraphs;
  sv.reqs = reqs;
  sv.nreqs = nreqs;
  sv.cfg = *cfg;
  IG_TRY(sv.arena.init(cfg->mem_budget));
  int hw_queues = 4;  // HIP's default hardware queues per process
  if (const char* q = std::getenv("GPU_MAX_HW_QUEUES")) {
    const int n = std::atoi(q);
    if (n > 0) hw_queues = n;
  }
  const int nslots = std::min(cfg->slots, hw_queues);
  sv.slots.resize(static_cast<size_t>(nslots));
  sv.stats.slots = nslots;
  int rc = INDIGO_OK;
  for (Slot& s : sv.slots) {
    rc = indigo_ctx_create(cfg->device, &s.ctx);
    if (rc != INDIGO_OK) break;
    // the GPU is shared: give persistent grids time to assemble
    s.ctx->hints[INDIGO_HINT_RESIDENCY_MS] = 2000;
  }
  // workspaces sized up front for the largest triangle count, so that no
  // slot reallocates (hipFree waits for the device) while others run
  size_t ws = 0;
  for (int i = 0; i < nreqs; i++)
  {
    const indigo_graph_ref& G = graphs[reqs[i].graph];
    if (reqs[i].alg == INDIGO_QUERY_TC) ws = std::max(ws, tc_workspace_bytes(G.nodes, G.edges));
    if (sv.single(reqs[i])) ws = std::max(ws, sv.single_workspace_bytes(reqs[i]));
  }
  for (Slot& s : sv.slots)
    if (rc == INDIGO_OK && s.ctx != nullptr) rc = ensure_workspace(s.ctx, std::max(ws, size_t(1) << 28));
  int limit = 0;
  if (rc == INDIGO_OK) rc = batch_grid_limit(sv.slots[0].ctx, &limit);
  sv.grid_per_slot = std::max(1, limit / nslots);
  // single-query persistent kernels take the slot's share of the GPU too: a
  // grid sized for the whole GPU cannot assemble beside the other slots
  for (Slot& s : sv.slots)
    if (s.ctx != nullptr) s.ctx->hints[INDIGO_HINT_MAX_WORKGROUPS] = sv.grid_per_slot;
  if (rc == INDIGO_OK) rc = sv.run();
  if (rc != INDIGO_OK) {
    const std::string msg = indigo_last_error();
    for (Slot& s : sv.slots)
      if (s.ctx != nullptr) (void)hipStreamSynchronize(s.ctx->stream);
    set_error(rc, "%s", msg.c_str());
  }
  for (Slot& s : sv.slots)
    if (s.ctx != nullptr) indigo_ctx_destroy(s.ctx);
  if (stats != nullptr) *stats = sv.stats;
  return rc;
}

}  // namespace indigo
