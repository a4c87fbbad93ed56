// host_checks.cpp -- drives the library's host code (file readers and
// writers, METIS text, conversion, generators, the initial-partition seams)
// and the CPU oracle on small inputs, for a build under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/san/Makefile).  Test infrastructure: any
// memory or UB error aborts the run; a failed comparison prints it and exits 1.
//
//   host_checks <scratch dir> <stand-in gpmetis dir>
#include "../../include/tgp.h"
#include "../../oracle/tgp_oracle.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static int failures = 0;
#define CHECK(cond)                                                           \
  do {                                                                        \
    if (!(cond)) {                                                            \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
      failures++;                                                             \
    }                                                                         \
  } while (0)

static std::string slurp(const std::string& path) {
  std::string s;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return s;
  char buf[4096];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, r);
  std::fclose(f);
  return s;
}

static void write_ints(const std::string& path, const int64_t* v, int64_t n) {
  FILE* f = std::fopen(path.c_str(), "w");
  for (int64_t i = 0; i < n; i++) std::fprintf(f, "%lld\n", (long long)v[i]);
  std::fclose(f);
}
static void write_ints32(const std::string& path, const int32_t* v, int64_t n) {
  FILE* f = std::fopen(path.c_str(), "w");
  for (int64_t i = 0; i < n; i++) std::fprintf(f, "%d\n", v[i]);
  std::fclose(f);
}

static bool same_graph(const tgp_graph& a, const tgp_graph& b) {
  if (a.n != b.n || a.m != b.m) return false;
  return !std::memcmp(a.xadj, b.xadj, sizeof(int64_t) * (a.n + 1)) &&
         !std::memcmp(a.adjncy, b.adjncy, sizeof(int32_t) * a.m) &&
         !std::memcmp(a.adjwgt, b.adjwgt, sizeof(int32_t) * a.m);
}

// METIS text, writers and readers, the oracle's partition, on one graph
static void check_graph(const tgp_graph& g, const std::string& dir, int32_t k) {
  // METIS text against the oracle's restatement
  char* text = nullptr;
  int64_t len = 0, np = 0;
  CHECK(tgp_metis_text(&g, &text, &len, &np) == TGP_OK);
  int64_t onp = 0;
  const int64_t olen =
      tgpo_metis_text(g.n, g.xadj, g.adjncy, g.adjwgt, nullptr, 0, &onp);
  std::vector<char> obuf((size_t)olen + 1);
  tgpo_metis_text(g.n, g.xadj, g.adjncy, g.adjwgt, obuf.data(), olen, &onp);
  CHECK(len == olen && np == onp && !std::memcmp(text, obuf.data(), (size_t)len));
  tgp_free(text);
  // reference directory form, README form and binary CSR: round trips
  write_ints(dir + "/row.txt", g.xadj, g.n + 1);
  write_ints32(dir + "/column.txt", g.adjncy, g.m);
  write_ints32(dir + "/weight.txt", g.adjwgt, g.m);
  tgp_graph r{};
  CHECK(tgp_read_graph_dir(dir.c_str(), g.n, g.m, &r) == TGP_OK);
  CHECK(same_graph(g, r));
  tgp_graph_free(&r);
  CHECK(tgp_read_graph_dir(dir.c_str(), g.n, g.m + 3, &r) != TGP_OK);  // short column.txt
  tgp_graph c{};
  CHECK(tgp_read_csr_files((dir + "/row.txt").c_str(), (dir + "/column.txt").c_str(), &c) ==
        TGP_OK);
  CHECK(c.n == g.n && c.m == g.m);
  tgp_graph_free(&c);
  const std::string bp = dir + "/g.bcsr";
  CHECK(tgp_write_bcsr(bp.c_str(), &g) == TGP_OK);
  tgp_graph b{};
  CHECK(tgp_read_bcsr(bp.c_str(), &b) == TGP_OK);
  CHECK(same_graph(g, b));
  tgp_graph_free(&b);
  // the oracle's partition and the writers of its result
  tgpo_options o{};
  o.k = k;
  o.repair = 1;
  o.coarsen_floor = 64;
  std::vector<int32_t> part((size_t)g.n), pc((size_t)k);
  int64_t cut = 0, tl = 0;
  std::vector<tgpo_trace> tr(4096);
  CHECK(tgpo_partition(g.n, g.xadj, g.adjncy, g.adjwgt, &o, part.data(), pc.data(), &cut,
                       tr.data(), (int64_t)tr.size(), &tl) == TGPO_OK);
  CHECK(cut == tgpo_edge_cut(g.n, g.xadj, g.adjncy, g.adjwgt, g.m, part.data(), k, 0));
  const std::string fp = dir + "/finalpartition.txt";
  CHECK(tgp_write_final_partition(fp.c_str(), part.data(), g.n, k) == TGP_OK);
  const int64_t flen = tgpo_final_text(g.n, part.data(), k, nullptr, 0);
  std::vector<char> fbuf((size_t)flen + 1);
  tgpo_final_text(g.n, part.data(), k, fbuf.data(), flen);
  CHECK(slurp(fp) == std::string(fbuf.data(), (size_t)flen));
  CHECK(tgp_write_part_file((dir + "/g.part").c_str(), part.data(), g.n) == TGP_OK);
  // reference semantics (may end in the fatal contraction exit)
  o.repair = 0;
  const int rc = tgpo_partition(g.n, g.xadj, g.adjncy, g.adjwgt, &o, part.data(), pc.data(),
                                &cut, tr.data(), (int64_t)tr.size(), &tl);
  CHECK(rc == TGPO_OK || rc == TGPO_EEDGES);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: host_checks <scratch dir> <stand-in gpmetis dir>\n");
    return 2;
  }
  const std::string dir = argv[1], stub = argv[2];
  tgp_graph g{};
  CHECK(tgp_gen_rmat(11, 8, 17, 1, &g) == TGP_OK);
  check_graph(g, dir, 8);
  // initial-partition seams on the same graph: block rule and gpmetis
  std::vector<int32_t> part((size_t)g.n);
  CHECK(tgp_initpart_block(&g, 0, 8, part.data(), nullptr) == TGP_OK);
  const std::string path = stub + ":" + (std::getenv("PATH") ? std::getenv("PATH") : "");
  setenv("PATH", path.c_str(), 1);
  const char* modes[] = {"", "short", "range", "fail", "nofile"};
  const int expect[] = {TGP_OK, TGP_OK, TGP_OK, TGP_EINITPART, TGP_EPARTFILE};
  for (int i = 0; i < 5; i++) {
    setenv("TGP_STUB_MODE", modes[i], 1);
    std::remove((dir + "/output.graph.part.8").c_str());  // no part file left from the last mode
    tgp_gpmetis_dir d{dir.c_str(), 0, 0};
    const int rc = tgp_initpart_gpmetis_dir(&g, 0, 8, part.data(), &d);
    if (rc != expect[i]) std::fprintf(stderr, "gpmetis stand-in mode '%s': %d\n", modes[i], rc);
    CHECK(rc == expect[i]);
  }
  tgp_graph_free(&g);
  CHECK(tgp_gen_grid3d(9, 7, 5, &g) == TGP_OK);
  check_graph(g, dir, 5);
  tgp_graph_free(&g);
  CHECK(tgp_gen_er(1500, 6.0, 3, &g) == TGP_OK);
  check_graph(g, dir, 4);
  tgp_graph_free(&g);
  CHECK(tgp_gen_rmat_n(2000, 8, 5, &g) == TGP_OK);
  check_graph(g, dir, 16);
  // conversion.py's inputs and outputs
  const std::string tsv = dir + "/edges.tsv";
  FILE* f = std::fopen(tsv.c_str(), "w");
  for (int i = 0; i < 200; i++) std::fprintf(f, "%d\t%d\t1\n", i, (i * 7 + 3) % 211);
  std::fclose(f);
  tgp_graph cv{};
  CHECK(tgp_convert_edges(tsv.c_str(), &cv) == TGP_OK);
  CHECK(tgp_write_conversion_files(&cv, dir.c_str()) == TGP_OK);
  tgp_graph_free(&cv);
  const std::string mtx = dir + "/edges.mtx";
  f = std::fopen(mtx.c_str(), "w");
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate pattern general\n50 50 60\n");
  for (int i = 0; i < 60; i++) std::fprintf(f, "%d %d\n", 1 + i % 50, 1 + (i * 13) % 50);
  std::fclose(f);
  CHECK(tgp_convert_edges(mtx.c_str(), &cv) == TGP_OK);
  tgp_graph_free(&cv);
  // error paths
  CHECK(tgp_read_bcsr((dir + "/missing.bcsr").c_str(), &cv) != TGP_OK);
  CHECK(tgp_read_graph_dir((dir + "/missing").c_str(), 3, 3, &cv) != TGP_OK);
  CHECK(std::strlen(tgp_strerror(TGP_EEDGES)) > 0);
  tgp_graph_free(&g);
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
