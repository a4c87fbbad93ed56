// gpu_part_main.cpp -- the `gpu_part` command, a drop-in for the reference
// program's process contract (src/gpu_part.cu:878-1390):
//
//   gpu_part <directory> <nvtxs> <nedges> <num_partitions> [options]
//
// reads <directory>/row.txt, column.txt, weight.txt; writes ./output.graph and
// runs `gpmetis -ptype=rb ./output.graph <k>` for the initial partition (the
// reference's seam), then ./finalpartition.txt and the same report lines.
// Errors print the reference's messages and exit with status 1.
//
// Two more input forms (no counterpart in the reference's code):
//   gpu_part <row.txt> <column.txt> <num_partitions> [options]   the README
//       form (README.md:133-141): sizes from the files, unit weights
//   gpu_part <graph.bcsr> <num_partitions> [options]             binary CSR
//
// Options after the positional arguments (none exist in the reference):
//   --init=block        deterministic i*k/n_c initial partition instead of gpmetis
//   --repair            opt-in repaired coarsening (tgp_options.repair)
//   --wide              64-bit cut and no n*k < 2^31 limit (tgp_options.wide)
//   --part-out=<path>   also write a METIS-style per-vertex part file
//   --write-bcsr=<path> also save the input graph as a binary CSR file
#include "../../include/tgp.h"

#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

static bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

int main(int argc, char* argv[]) {
  // which input form: binary CSR file, README row/column pair, or directory
  int form = 0;  // 0 directory, 1 README pair, 2 binary
  if (argc >= 3 && ends_with(argv[1], ".bcsr")) {
    form = 2;
  } else if (argc >= 4) {
    struct stat sb;
    if (stat(argv[1], &sb) == 0 && S_ISREG(sb.st_mode)) form = 1;
  }
  const int npos = form == 0 ? 4 : form == 1 ? 3 : 2;
  if (form == 0 && argc < 5) {  // gpu_part.cu:880-883
    std::cerr << "Usage: " << argv[0] << " <directory> <nvtxs> <nedges> <num_partitions>\n";
    return 1;
  }
  bool wide = false, repair = false, block = false;
  std::string part_out, bcsr_out;
  for (int i = npos + 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--wide")) wide = true;
    else if (!std::strcmp(argv[i], "--repair")) repair = true;
    else if (!std::strcmp(argv[i], "--init=block")) block = true;
    else if (!std::strcmp(argv[i], "--init=gpmetis")) block = false;
    else if (!std::strncmp(argv[i], "--part-out=", 11)) part_out = argv[i] + 11;
    else if (!std::strncmp(argv[i], "--write-bcsr=", 13)) bcsr_out = argv[i] + 13;
    else {
      std::cerr << "Unknown option: " << argv[i] << "\n";
      return 1;
    }
  }
  long long nv = 0, ne = 0;
  int k = 0;
  auto t0 = std::chrono::steady_clock::now();
  tgp_graph g;
  int rc;
  if (form == 0) {
    nv = wide ? std::atoll(argv[2]) : (long long)std::atoi(argv[2]);
    ne = wide ? std::atoll(argv[3]) : (long long)std::atoi(argv[3]);
    k = std::atoi(argv[4]);
    if (nv <= 0 || ne <= 0 || k <= 0) {  // gpu_part.cu:891-894
      std::cerr << "Invalid input: nvtxs, nedges, and numPartitions must be positive.\n";
      return 1;
    }
    rc = tgp_read_graph_dir(argv[1], nv, ne, &g);
  } else {
    k = std::atoi(argv[npos]);
    if (k <= 0) {
      std::cerr << "Invalid input: nvtxs, nedges, and numPartitions must be positive.\n";
      return 1;
    }
    rc = form == 1 ? tgp_read_csr_files(argv[1], argv[2], &g) : tgp_read_bcsr(argv[1], &g);
    if (rc == TGP_OK) {
      nv = g.n;
      ne = g.m;
    }
  }
  if (rc != TGP_OK) {
    std::cerr << tgp_last_error() << "\n";
    return 1;
  }
  if (!bcsr_out.empty() && tgp_write_bcsr(bcsr_out.c_str(), &g) != TGP_OK) {
    std::cerr << tgp_last_error() << "\n";
    tgp_graph_free(&g);
    return 1;
  }
  const long long target_vertices = std::max<long long>(500, nv / 500);  // gpu_part.cu:897

  const double t_read = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  tgp_options o;
  std::memset(&o, 0, sizeof(o));
  o.k = k;
  o.wide = wide;
  o.repair = repair;
  o.initpart = block ? tgp_initpart_block : tgp_initpart_gpmetis;
  o.profile = 1;
  tgp_stats st;
  std::memset(&st, 0, sizeof(st));
  std::vector<int32_t> part((size_t)nv), pc((size_t)k);
  int64_t cut = 0;
  rc = tgp_partition(&g, &o, part.data(), pc.data(), &cut, &st);
  if (rc != TGP_OK) {
    const char* msg = tgp_last_error();
    std::cerr << (msg && *msg ? msg : tgp_strerror(rc)) << "\n";
    tgp_graph_free(&g);
    return 1;
  }
  t0 = std::chrono::steady_clock::now();
  rc = tgp_write_final_partition("./finalpartition.txt", part.data(), nv, k);
  if (rc != TGP_OK) {
    std::cerr << "Error: Cannot open finalpartition.txt for writing\n";
    tgp_graph_free(&g);
    return 1;
  }
  if (!part_out.empty() && tgp_write_part_file(part_out.c_str(), part.data(), nv) != TGP_OK) {
    std::cerr << tgp_last_error() << "\n";
    tgp_graph_free(&g);
    return 1;
  }
  const double t_write = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // report (gpu_part.cu:1330-1344 for the result lines)
  std::cout << "Graph Detail ******************************************************\n";
  std::cout << std::left << std::setw(35) << "Number of Vertices:" << std::right << std::setw(10) << nv << "\n";
  std::cout << std::left << std::setw(35) << "Number of Edges:" << std::right << std::setw(10) << ne << "\n";
  std::cout << std::left << std::setw(35) << "Number of Partitions:" << std::right << std::setw(10) << k << "\n\n";
  std::cout << "K-way GPU with Deterministic Refinement Details *******************\n";
  std::cout << std::left << std::setw(35) << "Edge Cut:" << std::right << std::setw(10) << cut << "\n";
  std::cout << "Partition Sizes:\n";
  for (int q = 0; q < k; q++)
    std::cout << "  Partition " << q << ":" << std::right << std::setw(10) << pc[q] << " vertices\n";
  std::cout << std::left << std::setw(35) << "Balance Constraint:" << std::right << std::setw(10)
            << std::fixed << std::setprecision(1) << (0.03f * 100.0f) << "%\n";
  std::cout << std::left << std::setw(35) << "Coarsening Threshold:" << std::right << std::setw(10)
            << target_vertices << " vertices\n";
  std::cout << std::left << std::setw(35) << "Output File:" << std::right << std::setw(10)
            << "./finalpartition.txt" << "\n\n";

  double kernel_s = 0;
  for (int i = 0; i < TGP_NKERNELS; i++) kernel_s += st.kernel_ms[i] / 1000.0;
  std::cout << "Timing Breakdown **************************************************\n";
  std::cout << std::setprecision(6);
  std::cout << std::left << std::setw(35) << "  - File Reading Time:" << std::right << std::setw(10) << t_read << "s\n";
  std::cout << std::left << std::setw(35) << "  - Partition Time (wall):" << std::right << std::setw(10) << st.seconds_total << "s\n";
  std::cout << std::left << std::setw(35) << "  - Initial Partitioner Time:" << std::right << std::setw(10) << st.seconds_initpart << "s\n";
  std::cout << std::left << std::setw(35) << "  - GPU Kernel Time:" << std::right << std::setw(10) << kernel_s << "s\n";
  for (int i = 0; i < TGP_NKERNELS; i++) {
    const char* nm = tgp_kernel_name(i);
    if (!nm || !*nm || st.kernel_launches[i] == 0) continue;
    std::string label = std::string("      ") + nm + ":";
    std::cout << std::left << std::setw(35) << label << std::right << std::setw(10)
              << st.kernel_ms[i] / 1000.0 << "s\n";
  }
  std::cout << std::left << std::setw(35) << "  - Partition File Write Time:" << std::right << std::setw(10) << t_write << "s\n";
  std::cout << std::left << std::setw(35) << "Coarsening Levels:" << std::right << std::setw(10) << st.levels << "\n";
  std::cout << std::left << std::setw(35) << "Refinement Passes:" << std::right << std::setw(10) << st.passes << "\n\n";
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
    std::cout << "Memory Information ************************************************\n";
    std::cout << std::left << std::setw(35) << "Device Memory In Use:" << std::right << std::setw(10)
              << std::setprecision(2) << (double)(tot - fr) / (1024.0 * 1024.0) << " MB\n";
  }
  tgp_graph_free(&g);
  return 0;
}
