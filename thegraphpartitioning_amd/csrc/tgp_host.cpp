// tgp_host.cpp -- host side of the reference's file contract:
//   row.txt / column.txt / weight.txt readers (gpu_part.cu:786-800, 943-945),
//   the README form's row/column pair with unit weights (README.md:133-141),
//   a binary CSR file and a per-vertex part file (no counterpart in the
//   reference),
//   the METIS text of the coarsest graph and the gpmetis seam
//   (gpu_part.cu:803-864, 1157-1174), finalpartition.txt (gpu_part.cu:1294-1321).
#include "../../include/tgp.h"
#include "tgp_internal.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using tgp::i32;
using tgp::i64;
using tgp::u64;

namespace {

// Reads exactly `count` decimal ints the way fscanf("%d") does for
// well-formed input: skip white space, optional sign, digits.  Returns the
// index of the first element that could not be read, or count.  Row offsets
// are read into int64 (so row.txt can describe more than 2^31 arcs); ids and
// weights into int32.
template <typename T>
i64 read_ints(FILE* f, T* out, i64 count) {
  std::vector<char> buf(1 << 20);
  size_t len = 0, pos = 0;
  bool eof = false;
  auto peek = [&]() -> int {
    if (pos == len) {
      if (eof) return EOF;
      len = std::fread(buf.data(), 1, buf.size(), f);
      pos = 0;
      if (len == 0) {
        eof = true;
        return EOF;
      }
    }
    return (unsigned char)buf[pos];
  };
  for (i64 i = 0; i < count; i++) {
    int c = peek();
    while (c != EOF && (c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f')) {
      pos++;
      c = peek();
    }
    bool neg = false;
    if (c == '-' || c == '+') {
      neg = (c == '-');
      pos++;
      c = peek();
    }
    if (c == EOF || c < '0' || c > '9') return i;
    int64_t v = 0;
    while (c != EOF && c >= '0' && c <= '9') {
      v = v * 10 + (c - '0');
      if (v > (int64_t)1 << 62) v = (int64_t)1 << 62;
      pos++;
      c = peek();
    }
    out[i] = (T)(neg ? -v : v);
  }
  return count;
}

template <typename T>
int read_file(const std::string& path, T* out, i64 count) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    tgp::last_error() = "Error opening file: " + path + ": " + std::strerror(errno);
    return TGP_EIO;
  }
  const i64 got = read_ints(f, out, count);
  std::fclose(f);
  if (got != count) {
    tgp::last_error() = "Error reading file " + path + " at element " + std::to_string(got);
    return TGP_EIO;
  }
  return TGP_OK;
}

}  // namespace

extern "C" {

const char* tgp_strerror(int code) {
  switch (code) {
    case TGP_OK: return "success";
    case TGP_EINVAL: return "invalid argument or malformed CSR";
    case TGP_ELIMIT: return "reference int32 limit exceeded";
    case TGP_EEDGES: return "invalid coarse edge count";
    case TGP_EINITPART: return "initial partitioner failed";
    case TGP_EPARTFILE: return "cannot open the part file";
    case TGP_EHIP: return "HIP runtime error";
    case TGP_ENOMEM: return "out of memory";
    case TGP_EIO: return "file I/O error";
    case TGP_EDIFF: return "engine settings gave different results";
    default: return "unknown error";
  }
}

void tgp_free(void* p) { std::free(p); }

int tgp_read_graph_dir(const char* dir, int64_t n, int64_t m, tgp_graph* out) {
  if (!dir || !out || n <= 0 || m <= 0) return TGP_EINVAL;
  int64_t* xadj = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
  int32_t* adj = (int32_t*)std::malloc(sizeof(int32_t) * m);
  int32_t* w = (int32_t*)std::malloc(sizeof(int32_t) * m);
  if (!xadj || !adj || !w) {
    std::free(xadj);
    std::free(adj);
    std::free(w);
    return TGP_ENOMEM;
  }
  const std::string d(dir);
  int rc = read_file(d + "/row.txt", xadj, n + 1);
  if (rc == TGP_OK) rc = read_file(d + "/column.txt", adj, m);
  if (rc == TGP_OK) rc = read_file(d + "/weight.txt", w, m);
  if (rc != TGP_OK) {
    std::free(xadj);
    std::free(adj);
    std::free(w);
    return rc;
  }
  out->n = n;
  out->m = m;
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  return TGP_OK;
}

// Counts the whitespace-separated decimal ints of a file (stops at the first
// token that is not one, as the reference's fscanf loop would).
static int count_ints(const std::string& path, i64* count) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    tgp::last_error() = "Error opening file: " + path + ": " + std::strerror(errno);
    return TGP_EIO;
  }
  std::vector<char> buf(1 << 20);
  i64 c = 0;
  bool in = false, ok = true;
  size_t len;
  while (ok && (len = std::fread(buf.data(), 1, buf.size(), f)) > 0) {
    for (size_t i = 0; i < len; i++) {
      const char ch = buf[i];
      const bool digit = (ch >= '0' && ch <= '9') || ch == '-' || ch == '+';
      const bool space = ch == ' ' || ch == '\n' || ch == '\t' || ch == '\r' || ch == '\v' ||
                         ch == '\f';
      if (digit) {
        if (!in) c++;
        in = true;
      } else if (space) {
        in = false;
      } else {
        ok = false;
        break;
      }
    }
  }
  std::fclose(f);
  *count = c;
  return TGP_OK;
}

int tgp_read_csr_files(const char* row_path, const char* col_path, tgp_graph* out) {
  if (!row_path || !col_path || !out) return TGP_EINVAL;
  i64 nr = 0, m = 0;
  int rc = count_ints(row_path, &nr);
  if (rc == TGP_OK) rc = count_ints(col_path, &m);
  if (rc != TGP_OK) return rc;
  const i64 n = nr - 1;
  if (n <= 0 || m <= 0) {
    tgp::last_error() = "empty row or column file";
    return TGP_EINVAL;
  }
  int64_t* xadj = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
  int32_t* adj = (int32_t*)std::malloc(sizeof(int32_t) * m);
  int32_t* w = (int32_t*)std::malloc(sizeof(int32_t) * m);
  if (!xadj || !adj || !w) {
    std::free(xadj);
    std::free(adj);
    std::free(w);
    return TGP_ENOMEM;
  }
  rc = read_file(row_path, xadj, nr);
  if (rc == TGP_OK) rc = read_file(col_path, adj, m);
  if (rc != TGP_OK) {
    std::free(xadj);
    std::free(adj);
    std::free(w);
    return rc;
  }
  for (i64 i = 0; i < m; i++) w[i] = 1;
  out->n = n;
  out->m = m;
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  return TGP_OK;
}

// Binary CSR: "TGPCSR01", int64 n, int64 m, int64 xadj[n+1], int32 adj[m],
// int32 w[m], little-endian.
static const char BCSR_MAGIC[8] = {'T', 'G', 'P', 'C', 'S', 'R', '0', '1'};

int tgp_write_bcsr(const char* path, const tgp_graph* g) {
  if (!path || !g || g->n <= 0 || g->m < 0) return TGP_EINVAL;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    tgp::last_error() = std::string("Error opening file: ") + path + ": " + std::strerror(errno);
    return TGP_EIO;
  }
  const int64_t hdr[2] = {g->n, g->m};
  bool ok = std::fwrite(BCSR_MAGIC, 1, 8, f) == 8 && std::fwrite(hdr, 8, 2, f) == 2 &&
            std::fwrite(g->xadj, 8, (size_t)g->n + 1, f) == (size_t)g->n + 1 &&
            std::fwrite(g->adjncy, 4, (size_t)g->m, f) == (size_t)g->m &&
            std::fwrite(g->adjwgt, 4, (size_t)g->m, f) == (size_t)g->m;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? TGP_OK : TGP_EIO;
}

int tgp_read_bcsr(const char* path, tgp_graph* out) {
  if (!path || !out) return TGP_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    tgp::last_error() = std::string("Error opening file: ") + path + ": " + std::strerror(errno);
    return TGP_EIO;
  }
  char magic[8];
  int64_t hdr[2] = {0, 0};
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, BCSR_MAGIC, 8) != 0 ||
      std::fread(hdr, 8, 2, f) != 2 || hdr[0] <= 0 || hdr[1] < 0) {
    std::fclose(f);
    tgp::last_error() = std::string("not a binary CSR file: ") + path;
    return TGP_EIO;
  }
  const int64_t n = hdr[0], m = hdr[1];
  int64_t* xadj = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
  int32_t* adj = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(m, 1));
  int32_t* w = (int32_t*)std::malloc(sizeof(int32_t) * std::max<int64_t>(m, 1));
  const bool ok = xadj && adj && w &&
                  std::fread(xadj, 8, (size_t)n + 1, f) == (size_t)n + 1 &&
                  std::fread(adj, 4, (size_t)m, f) == (size_t)m &&
                  std::fread(w, 4, (size_t)m, f) == (size_t)m;
  std::fclose(f);
  if (!ok) {
    std::free(xadj);
    std::free(adj);
    std::free(w);
    tgp::last_error() = std::string("truncated binary CSR file: ") + path;
    return TGP_EIO;
  }
  out->n = n;
  out->m = m;
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  return TGP_OK;
}

int tgp_write_part_file(const char* path, const int32_t* part, int64_t n) {
  if (!path || !part || n <= 0) return TGP_EINVAL;
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    tgp::last_error() = std::string("Error opening file: ") + path + ": " + std::strerror(errno);
    return TGP_EIO;
  }
  std::string buf;
  char num[16];
  for (i64 i = 0; i < n; i++) {
    const int len = std::snprintf(num, sizeof(num), "%d\n", part[i]);
    buf.append(num, (size_t)len);
    if (buf.size() > (1 << 20)) {
      std::fwrite(buf.data(), 1, buf.size(), f);
      buf.clear();
    }
  }
  std::fwrite(buf.data(), 1, buf.size(), f);
  return std::fclose(f) == 0 ? TGP_OK : TGP_EIO;
}

int tgp_write_final_partition(const char* path, const int32_t* part, int64_t n, int32_t k) {
  if (!path || !part || n <= 0 || k <= 0) return TGP_EINVAL;
  // vertices grouped by part, ascending; out-of-range parts go to part 0
  std::vector<i64> cnt((size_t)k + 1, 0);
  for (i64 i = 0; i < n; i++) {
    const i32 p = (part[i] >= 0 && part[i] < k) ? part[i] : 0;
    cnt[p + 1]++;
  }
  for (i32 q = 0; q < k; q++) cnt[q + 1] += cnt[q];
  std::vector<i64> order((size_t)n);
  {
    std::vector<i64> fill(cnt.begin(), cnt.end() - 1);
    for (i64 i = 0; i < n; i++) {
      const i32 p = (part[i] >= 0 && part[i] < k) ? part[i] : 0;
      order[fill[p]++] = i;
    }
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    tgp::last_error() = "Error: Cannot open finalpartition.txt for writing";
    return TGP_EIO;
  }
  std::string line;
  char num[32];
  for (i32 q = 0; q < k; q++) {
    line.clear();
    line += "Partition ";
    line += std::to_string(q);
    line += ": ";
    for (i64 j = cnt[q]; j < cnt[q + 1]; j++) {
      const int len = std::snprintf(num, sizeof(num), "%lld", (long long)order[j]);
      line.append(num, (size_t)len);
      if (j + 1 < cnt[q + 1]) line += ',';
    }
    line += '\n';
    std::fwrite(line.data(), 1, line.size(), f);
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? TGP_OK : TGP_EIO;
}

int tgp_metis_text(const tgp_graph* g, char** text, int64_t* len, int64_t* npairs) {
  if (!g || !text || !len || !npairs || g->n <= 0) return TGP_EINVAL;
  const i64 n = g->n;
  // canonical pairs (min,max) from every arc with 0 <= u < n, u != i over the
  // true row ranges; weight = sum over both directions (int32 wrap)
  std::vector<u64> keys;
  std::vector<i32> wts;
  keys.reserve((size_t)g->m);
  wts.reserve((size_t)g->m);
  for (i64 i = 0; i < n; i++) {
    for (i64 j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
      const i32 u = g->adjncy[j];
      if (u >= 0 && u < n && u != i) {
        const u64 a = (u64)std::min<i64>(i, u), b = (u64)std::max<i64>(i, u);
        keys.push_back((a << 32) | b);
        wts.push_back(g->adjwgt[j]);
      }
    }
  }
  std::vector<i64> idx(keys.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = (i64)i;
  std::sort(idx.begin(), idx.end(), [&](i64 a, i64 b) { return keys[a] < keys[b]; });
  std::vector<u64> pk;
  std::vector<i32> pw;
  for (i64 t : idx) {
    if (!pk.empty() && pk.back() == keys[t])
      pw.back() = (i32)((uint32_t)pw.back() + (uint32_t)wts[t]);
    else {
      pk.push_back(keys[t]);
      pw.push_back(wts[t]);
    }
  }
  *npairs = (int64_t)pk.size();
  // adjacency lists: for line i the neighbours below i (pairs (u,i)) come in
  // ascending u, then those above i (pairs (i,v)) in ascending v
  std::vector<i64> deg((size_t)n + 1, 0);
  for (u64 k2 : pk) {
    deg[(k2 >> 32) + 1]++;
    deg[(k2 & 0xffffffffu) + 1]++;
  }
  for (i64 i = 0; i < n; i++) deg[i + 1] += deg[i];
  std::vector<i32> nb((size_t)deg[n]), nw((size_t)deg[n]);
  {
    std::vector<i64> fill(deg.begin(), deg.end() - 1);
    for (size_t p = 0; p < pk.size(); p++) {
      const i64 a = (i64)(pk[p] >> 32), b = (i64)(pk[p] & 0xffffffffu);
      nb[fill[a]] = (i32)b;
      nw[fill[a]++] = pw[p];
      nb[fill[b]] = (i32)a;
      nw[fill[b]++] = pw[p];
    }
  }
  std::string s = std::to_string(n) + " " + std::to_string(pk.size()) + " 001\n";
  char num[32];
  for (i64 i = 0; i < n; i++) {
    // pairs were filled in (min,max) order, so entries (u,i) with u < i precede
    // entries (i,v): the line is already ascending
    for (i64 j = deg[i]; j < deg[i + 1]; j++) {
      if (j > deg[i]) s += ' ';
      int l = std::snprintf(num, sizeof(num), "%lld %d", (long long)nb[j] + 1, nw[j]);
      s.append(num, (size_t)l);
    }
    s += '\n';
  }
  char* out = (char*)std::malloc(s.size() + 1);
  if (!out) return TGP_ENOMEM;
  std::memcpy(out, s.data(), s.size());
  out[s.size()] = 0;
  *text = out;
  *len = (int64_t)s.size();
  return TGP_OK;
}

// Writes <dir>output.graph, runs gpmetis on it and reads the part file
// (gpu_part.cu:829-864, 1157-1174).  An empty `dir` is the working directory
// with the reference's own names and command; otherwise the command quotes
// the graph path.
static int gpmetis_in(const std::string& dir, const tgp_graph* coarsest, int32_t k,
                      int32_t* part, int* status) {
  char* text = nullptr;
  int64_t len = 0, np = 0;
  int rc = tgp_metis_text(coarsest, &text, &len, &np);
  if (rc != TGP_OK) return rc;
  const std::string gpath = dir + "output.graph";
  FILE* f = std::fopen(gpath.c_str(), "wb");
  if (!f) {
    std::free(text);
    tgp::last_error() = "Error: Cannot create " + gpath;
    return TGP_EIO;
  }
  std::fwrite(text, 1, (size_t)len, f);
  std::fclose(f);
  std::free(text);
  const std::string cmd = "gpmetis -ptype=rb " +
                          (dir.empty() ? std::string("./output.graph") : "'" + gpath + "'") +
                          " " + std::to_string(k);
  const int ret = std::system(cmd.c_str());
  if (status) *status = ret;
  if (ret != 0) {
    tgp::last_error() = "Error: gpmetis failed with return code " + std::to_string(ret);
    return TGP_EINITPART;
  }
  const std::string pf = gpath + ".part." + std::to_string(k);
  std::ifstream fin(pf);
  if (!fin.is_open()) {
    tgp::last_error() = "Error: Cannot open " + pf;
    return TGP_EPARTFILE;
  }
  std::vector<int> h((size_t)coarsest->n, 0);
  for (i64 i = 0; i < coarsest->n; i++) {
    fin >> h[i];
    if (h[i] < 0 || h[i] >= k) h[i] = 0;
  }
  for (i64 i = 0; i < coarsest->n; i++) part[i] = h[i];
  return TGP_OK;
}

int tgp_initpart_gpmetis(const tgp_graph* coarsest, int64_t npairs, int32_t k, int32_t* part,
                         void* user) {
  (void)npairs;
  return gpmetis_in("", coarsest, k, part, (int*)user);
}

int tgp_initpart_gpmetis_dir(const tgp_graph* coarsest, int64_t npairs, int32_t k,
                             int32_t* part, void* user) {
  (void)npairs;
  tgp_gpmetis_dir* d = (tgp_gpmetis_dir*)user;
  if (!d || !d->dir || !*d->dir || std::strchr(d->dir, '\'')) {
    tgp::last_error() = "Error: gpmetis working directory missing or holds a quote";
    return TGP_EINVAL;
  }
  std::string dir(d->dir);
  if (dir.back() != '/') dir += '/';
  return gpmetis_in(dir, coarsest, k, part, &d->status);
}

int tgp_initpart_block(const tgp_graph* coarsest, int64_t npairs, int32_t k, int32_t* part,
                       void* user) {
  (void)npairs;
  (void)user;
  const i64 n = coarsest->n;
  for (i64 i = 0; i < n; i++) part[i] = (i32)((i * (i64)k) / n);
  return TGP_OK;
}

}  // extern "C"
