// Host-code check of the CSR builders, generators and .egr I/O
// (parallelization-artifact-sc23_amd/csrc/graphgen.cpp), compiled with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_asan.py.
// Every case prints one line; any sanitizer report, leak or failed check
// makes the process exit non-zero.
//
// Usage: asan_graph_host <scratch directory>
#include "indigo_graph.h"
#include "indigo_hip.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int failures = 0;

void expect(bool ok, const char* what)
{
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) failures++;
}

bool same_graph(const indigo_host_graph& a, const indigo_host_graph& b)
{
  if (a.nodes != b.nodes || a.edges != b.edges) return false;
  if ((a.eweight == nullptr) != (b.eweight == nullptr)) return false;
  if (std::memcmp(a.nindex, b.nindex, sizeof(int) * (a.nodes + 1)) != 0) return false;
  if (std::memcmp(a.nlist, b.nlist, sizeof(int) * a.edges) != 0) return false;
  return a.eweight == nullptr || std::memcmp(a.eweight, b.eweight, sizeof(int) * a.edges) == 0;
}

void write_ints(const std::string& path, const std::vector<int>& v)
{
  FILE* f = std::fopen(path.c_str(), "wb");
  if (f == nullptr) std::abort();
  if (!v.empty() && std::fwrite(v.data(), sizeof(int), v.size(), f) != v.size()) std::abort();
  std::fclose(f);
}

// Writes the ints, then checks that both readers give `want` (INDIGO_OK, or
// the error status) and, on success, the same arrays.
void egr_case(const std::string& dir, const char* name, const std::vector<int>& ints, bool want_ok,
              bool want_weights)
{
  const std::string path = dir + "/" + name + ".egr";
  write_ints(path, ints);
  indigo_host_graph g{};
  const int rc = indigo_read_egr(path.c_str(), &g);
  indigo_egr_map m{};
  const int rm = indigo_egr_open(path.c_str(), &m);
  std::string what = std::string("egr ") + name;
  expect((rc == INDIGO_OK) == want_ok && (rm == INDIGO_OK) == want_ok, what.c_str());
  if (want_ok && rc == INDIGO_OK && rm == INDIGO_OK) {
    bool same = g.nodes == m.nodes && g.edges == m.edges &&
                std::memcmp(g.nindex, m.nindex, sizeof(int) * (g.nodes + 1)) == 0 &&
                (g.edges == 0 || std::memcmp(g.nlist, m.nlist, sizeof(int) * g.edges) == 0) &&
                (g.eweight != nullptr) == want_weights && (m.eweight != nullptr) == want_weights &&
                (!want_weights || std::memcmp(g.eweight, m.eweight, sizeof(int) * g.edges) == 0);
    expect(same, (what + " readers agree").c_str());
  } else {
    expect(g.nindex == nullptr && g.nlist == nullptr && g.eweight == nullptr, (what + " leaves nothing").c_str());
  }
  indigo_graph_release(&g);
  indigo_egr_close(&m);
}

}  // namespace

int main(int argc, char** argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];

  // generators: simple, sorted, symmetric, weights symmetric; thread count
  // does not change the result
  indigo_host_graph g1{}, g2{};
  expect(indigo_gen_grid(17, 23, 3, 1, 9, &g1) == INDIGO_OK && indigo_graph_check(&g1) == 0, "grid 17x23");
  indigo_graph_release(&g1);
  expect(indigo_gen_rmat(10, 8, .57, .19, .19, 1, 1, 255, 1, &g1) == INDIGO_OK && indigo_graph_check(&g1) == 0,
         "rmat 10 one thread");
  expect(indigo_gen_rmat(10, 8, .57, .19, .19, 1, 1, 255, 7, &g2) == INDIGO_OK && same_graph(g1, g2),
         "rmat 10 seven threads, same graph");
  indigo_graph_release(&g2);
  expect(indigo_gen_er(1000, 4000, 5, 0, -1, 3, &g2) == INDIGO_OK && indigo_graph_check(&g2) == 0 &&
             g2.eweight == nullptr,
         "er 1000 without weights");
  indigo_graph_release(&g2);
  expect(indigo_gen_road(31, 29, .85, .02, 9, 1, 1000, 4, &g2) == INDIGO_OK && indigo_graph_check(&g2) == 0,
         "road 31x29");
  indigo_graph_release(&g2);

  // build_csr: duplicates, self loops, out-of-range ids, empty input
  {
    const int src[] = {0, 1, 1, 2, 3, 3};
    const int dst[] = {1, 0, 1, 3, 2, 0};
    expect(indigo_build_csr(5, 6, src, dst, 1, 0, 0, -1, 2, &g2) == INDIGO_OK && indigo_graph_check(&g2) == 0 &&
               g2.edges == 6,
           "build_csr merges duplicates and drops self loops");
    indigo_graph_release(&g2);
    const int bad[] = {0, 5};
    expect(indigo_build_csr(5, 2, bad, bad, 1, 0, 0, -1, 2, &g2) == INDIGO_ERR_INVALID_ARG && g2.nindex == nullptr,
           "build_csr rejects an id out of range");
    expect(indigo_build_csr(4, 0, nullptr, nullptr, 1, 0, 0, -1, 1, &g2) == INDIGO_OK && g2.edges == 0 &&
               indigo_graph_check(&g2) == 0,
           "build_csr of no pairs");
    indigo_graph_release(&g2);
  }

  // graph_check names the first violation, also for missing arrays
  {
    indigo_host_graph h{};
    expect(indigo_graph_copy(&g1, &h) == INDIGO_OK && same_graph(g1, h), "copy");
    const int v = h.nlist[h.nindex[0]];
    h.nlist[h.nindex[0]] = h.nodes;
    expect(indigo_graph_check(&h) == 2, "check: id out of range");
    h.nlist[h.nindex[0]] = 0;
    expect(indigo_graph_check(&h) == 4 || indigo_graph_check(&h) == 3, "check: self loop");
    h.nlist[h.nindex[0]] = v;
    h.eweight[0] += 1;
    expect(indigo_graph_check(&h) == 6, "check: asymmetric weight");
    int* nl = h.nlist;
    h.nlist = nullptr;
    expect(indigo_graph_check(&h) == 1, "check: no nlist");
    h.nlist = nl;
    indigo_graph_release(&h);
  }

  // .egr round trip, then damaged files through both readers
  {
    const std::string path = dir + "/rmat10.egr";
    indigo_host_graph r{};
    expect(indigo_write_egr(path.c_str(), &g1) == INDIGO_OK && indigo_read_egr(path.c_str(), &r) == INDIGO_OK &&
               same_graph(g1, r),
           "egr round trip");
    indigo_graph_release(&r);
  }
  // path 0-1-2: header, nindex, nlist, weights
  const std::vector<int> head = {3, 4, 0, 1, 3, 4, 1, 0, 2, 1};
  const std::vector<int> wts = {5, 5, 7, 7};
  std::vector<int> full = head;
  full.insert(full.end(), wts.begin(), wts.end());
  egr_case(dir, "weighted", full, true, true);
  egr_case(dir, "unweighted", head, true, false);
  std::vector<int> extra = full;
  extra.push_back(99);
  egr_case(dir, "trailing_bytes", extra, true, true);
  egr_case(dir, "partial_weights", std::vector<int>(full.begin(), full.end() - 1), false, false);
  egr_case(dir, "truncated_nlist", std::vector<int>(head.begin(), head.end() - 1), false, false);
  egr_case(dir, "truncated_nindex", std::vector<int>(head.begin(), head.begin() + 4), false, false);
  egr_case(dir, "header_only", std::vector<int>(head.begin(), head.begin() + 2), false, false);
  egr_case(dir, "short_header", std::vector<int>(head.begin(), head.begin() + 1), false, false);
  egr_case(dir, "empty", {}, false, false);
  egr_case(dir, "zero_nodes", {0, 0, 0}, false, false);
  egr_case(dir, "negative_edges", {2, -1, 0, 0, 0}, false, false);
  egr_case(dir, "no_edges", {2, 0, 0, 0, 0}, true, false);
  {
    indigo_host_graph r{};
    indigo_egr_map m{};
    const std::string missing = dir + "/does_not_exist.egr";
    expect(indigo_read_egr(missing.c_str(), &r) == INDIGO_ERR_INVALID_ARG &&
               indigo_egr_open(missing.c_str(), &m) == INDIGO_ERR_INVALID_ARG,
           "egr missing file");
    expect(indigo_read_egr(nullptr, &r) == INDIGO_ERR_INVALID_ARG && indigo_write_egr(nullptr, &g1) ==
               INDIGO_ERR_INVALID_ARG,
           "egr null path");
  }
  indigo_graph_release(&g1);
  expect(indigo_graph_release(nullptr) == INDIGO_ERR_INVALID_ARG && indigo_egr_close(nullptr) == INDIGO_OK,
         "release of no graph is an error, close of no map is not");

  std::printf("%d failure(s)\n", failures);
  return failures == 0 ? 0 : 1;
}
