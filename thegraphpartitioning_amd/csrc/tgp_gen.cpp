// tgp_gen.cpp -- seeded synthetic CSR graphs for the partitioner's workloads.
//
// Inputs are produced in the same form the reference's preprocessing emits
// (preprocess/conversion.py:66-93): self-loops removed, every edge present in
// both directions, isolated vertices removed with an order-preserving relabel,
// rows listing their neighbours in ascending order.
//
//   R-MAT(scale s, edge factor ef, a,b,c,d = 0.57,0.19,0.19,0.05): ef*2^s
//     samples, no noise, no vertex scrambling.  R-MAT on n vertices keeps the
//     first ef*n samples of the scale ceil(log2 n) model inside [0, n)^2.
//   3-D 7-point grid nx*ny*nz: id = x + nx*y + nx*ny*z.
//   Erdos-Renyi G(n, p) by geometric skip sampling (Batagelj-Brandes).
//
// Random numbers come from a counter-based SplitMix64 keyed by
// (seed, stream, index), so the same graph results on every machine and
// thread count.
#include "../../include/tgp.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull + 0x9E3779B97F4A7C15ull) ^
               (idx + 1) * 0x9E3779B97F4A7C15ull);
}

inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// Symmetric per-edge weight 1 + H(seed, min, max) mod 16.
inline int32_t edge_weight(uint64_t seed, int32_t u, int32_t v) {
  const uint64_t a = (uint64_t)(uint32_t)std::min(u, v), b = (uint64_t)(uint32_t)std::max(u, v);
  return 1 + (int32_t)(draw(seed, 7, (a << 32) | b) % 16ull);
}

// Builds the symmetric, deduplicated, isolated-free CSR from an arc list
// (both directions already present).  Arcs are (src,dst) pairs, self-loops
// already dropped.  Returns a malloc'd graph.
int build_csr(int64_t nv, std::vector<int32_t>& src, std::vector<int32_t>& dst,
              int weight_mode, uint64_t seed, tgp_graph* out) {
  const int64_t na = (int64_t)src.size();
  std::vector<int64_t> deg((size_t)nv + 1, 0);
  {
    std::vector<std::atomic<int64_t>> cnt((size_t)nv);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; v++) cnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < na; i++) cnt[src[i]].fetch_add(1, std::memory_order_relaxed);
    for (int64_t v = 0; v < nv; v++) deg[v + 1] = deg[v] + cnt[v].load();
  }
  std::vector<int32_t> col((size_t)na);
  {
    std::vector<std::atomic<int64_t>> fill((size_t)nv);
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nv; v++) fill[v].store(deg[v], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < na; i++) {
      const int64_t p = fill[src[i]].fetch_add(1, std::memory_order_relaxed);
      col[p] = dst[i];
    }
  }
  std::vector<int32_t>().swap(src);
  std::vector<int32_t>().swap(dst);
  std::vector<int64_t> ulen((size_t)nv, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < nv; v++) {
    int32_t* b = col.data() + deg[v];
    int32_t* e = col.data() + deg[v + 1];
    std::sort(b, e);
    ulen[v] = std::unique(b, e) - b;
  }
  // order-preserving relabel of non-isolated vertices
  std::vector<int32_t> nid((size_t)nv, -1);
  int64_t n = 0;
  for (int64_t v = 0; v < nv; v++)
    if (ulen[v] > 0) nid[v] = (int32_t)n++;
  if (n == 0) return TGP_EINVAL;
  int64_t* xadj = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
  xadj[0] = 0;
  {
    int64_t i = 0;
    for (int64_t v = 0; v < nv; v++)
      if (ulen[v] > 0) { xadj[i + 1] = xadj[i] + ulen[v]; i++; }
  }
  const int64_t m = xadj[n];
  int32_t* adj = (int32_t*)std::malloc(sizeof(int32_t) * m);
  int32_t* w = (int32_t*)std::malloc(sizeof(int32_t) * m);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < nv; v++) {
    if (ulen[v] == 0) continue;
    const int64_t o = xadj[nid[v]];
    const int32_t u = nid[v];
    for (int64_t j = 0; j < ulen[v]; j++) {
      const int32_t x = nid[col[deg[v] + j]];
      adj[o + j] = x;
      w[o + j] = weight_mode == 1 ? edge_weight(seed, u, x) : 1;
    }
  }
  out->n = n;
  out->m = m;
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  return TGP_OK;
}

}  // namespace

extern "C" {

int tgp_gen_rmat(int32_t scale, int32_t edge_factor, uint64_t seed, int32_t weight_mode,
                 tgp_graph* out) {
  if (!out || scale < 1 || scale > 30 || edge_factor < 1) return TGP_EINVAL;
  const int64_t nv = 1ll << scale;
  const int64_t ns = nv * edge_factor;
  const double a = 0.57, b = 0.19, c = 0.19;
  std::vector<int32_t> su((size_t)ns), sv((size_t)ns);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ns; i++) {
    int64_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      const double r = unit(draw(seed, (uint64_t)l + 1, (uint64_t)i));
      int bu, bv;
      if (r < a) { bu = 0; bv = 0; }
      else if (r < a + b) { bu = 0; bv = 1; }
      else if (r < a + b + c) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    su[i] = (int32_t)u;
    sv[i] = (int32_t)v;
  }
  // drop self-loops, add reverse arcs
  std::vector<int64_t> keep((size_t)ns);
  int64_t na = 0;
  for (int64_t i = 0; i < ns; i++) {
    keep[i] = na;
    if (su[i] != sv[i]) na += 2;
  }
  std::vector<int32_t> src((size_t)na), dst((size_t)na);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ns; i++) {
    if (su[i] == sv[i]) continue;
    const int64_t p = keep[i];
    src[p] = su[i]; dst[p] = sv[i];
    src[p + 1] = sv[i]; dst[p + 1] = su[i];
  }
  std::vector<int32_t>().swap(su);
  std::vector<int32_t>().swap(sv);
  std::vector<int64_t>().swap(keep);
  return build_csr(nv, src, dst, weight_mode, seed, out);
}

int tgp_gen_rmat_n(int64_t n, int32_t edge_factor, uint64_t seed, tgp_graph* out) {
  if (!out || n < 2 || n > INT32_MAX || edge_factor < 1) return TGP_EINVAL;
  int scale = 1;
  while (((int64_t)1 << scale) < n) scale++;
  const int64_t want = n * edge_factor;
  const double a = 0.57, b = 0.19, c = 0.19;
  std::vector<int32_t> src, dst;
  src.reserve((size_t)want * 2);
  dst.reserve((size_t)want * 2);
  int64_t kept = 0;
  for (uint64_t i = 0; kept < want; i++) {
    int64_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      const double r = unit(draw(seed, (uint64_t)l + 1, i));
      const int bu = r >= a + b;
      const int bv = (r >= a && r < a + b) || r >= a + b + c;
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u >= n || v >= n) continue;
    kept++;
    if (u == v) continue;
    src.push_back((int32_t)u); dst.push_back((int32_t)v);
    src.push_back((int32_t)v); dst.push_back((int32_t)u);
  }
  return build_csr(n, src, dst, 0, seed, out);
}

int tgp_gen_grid3d(int32_t nx, int32_t ny, int32_t nz, tgp_graph* out) {
  if (!out || nx < 1 || ny < 1 || nz < 1) return TGP_EINVAL;
  const int64_t n = (int64_t)nx * ny * nz;
  if (n > INT32_MAX) return TGP_EINVAL;
  int64_t* xadj = (int64_t*)std::malloc(sizeof(int64_t) * (n + 1));
  auto degree = [&](int64_t id) {
    const int64_t x = id % nx, y = (id / nx) % ny, z = id / ((int64_t)nx * ny);
    return (int64_t)(x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny) + (z > 0) + (z + 1 < nz);
  };
  xadj[0] = 0;
  for (int64_t i = 0; i < n; i++) xadj[i + 1] = xadj[i] + degree(i);
  const int64_t m = xadj[n];
  if (m == 0) { std::free(xadj); return TGP_EINVAL; }
  int32_t* adj = (int32_t*)std::malloc(sizeof(int32_t) * m);
  int32_t* w = (int32_t*)std::malloc(sizeof(int32_t) * m);
  const int64_t sxy = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const int64_t x = i % nx, y = (i / nx) % ny, z = i / sxy;
    int64_t p = xadj[i];
    if (z > 0) adj[p++] = (int32_t)(i - sxy);
    if (y > 0) adj[p++] = (int32_t)(i - nx);
    if (x > 0) adj[p++] = (int32_t)(i - 1);
    if (x + 1 < nx) adj[p++] = (int32_t)(i + 1);
    if (y + 1 < ny) adj[p++] = (int32_t)(i + nx);
    if (z + 1 < nz) adj[p++] = (int32_t)(i + sxy);
    for (int64_t j = xadj[i]; j < p; j++) w[j] = 1;
  }
  out->n = n;
  out->m = m;
  out->xadj = xadj;
  out->adjncy = adj;
  out->adjwgt = w;
  return TGP_OK;
}

int tgp_gen_er(int64_t n, double mean_degree, uint64_t seed, tgp_graph* out) {
  if (!out || n < 2 || n > INT32_MAX || !(mean_degree > 0)) return TGP_EINVAL;
  const double p = std::min(1.0, mean_degree / (double)(n - 1));
  // Batagelj-Brandes over the strict lower triangle (v > u), one stream per row
  // block so the result does not depend on the thread count.
  std::vector<int32_t> src, dst;
  const double lq = std::log(1.0 - p);
  int64_t v = 1, u = -1, i = 0;
  while (v < n) {
    const double r = unit(draw(seed, 11, (uint64_t)i++));
    u += 1 + (p >= 1.0 ? 0 : (int64_t)std::floor(std::log(1.0 - r) / lq));
    while (u >= v && v < n) { u -= v; v++; }
    if (v < n) {
      src.push_back((int32_t)v); dst.push_back((int32_t)u);
      src.push_back((int32_t)u); dst.push_back((int32_t)v);
    }
  }
  return build_csr(n, src, dst, 0, seed, out);
}

// ---- edge-list files in the forms preprocess/conversion.py reads (30-58):
// .mtx/.mmio (Matrix Market coordinate, 1-based; the graph's arcs run from
// column to row, conversion.py:33-37), .tsv (source, dest, data), .txt (four
// header lines skipped), .tsv_1 (one header line skipped).  Then 66-69:
// self loops dropped, every edge in both directions once, isolated vertices
// removed with an order-preserving relabel; unit weights.  Rows list their
// neighbours in ascending order (DGL's own order is not specified).
int tgp_convert_edges(const char* path, tgp_graph* out) {
  if (!path || !out) return TGP_EINVAL;
  const std::string p(path);
  const size_t dot = p.find_last_of('.');
  const std::string ext = dot == std::string::npos ? "" : p.substr(dot);
  int skip = 0;
  bool mtx = false;
  if (ext == ".mtx" || ext == ".mmio") mtx = true;
  else if (ext == ".tsv") skip = 0;
  else if (ext == ".txt") skip = 4;
  else if (ext == ".tsv_1") skip = 1;
  else return TGP_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) return TGP_EIO;
  std::vector<char> text;
  {
    std::vector<char> buf(1 << 22);
    size_t len;
    while ((len = std::fread(buf.data(), 1, buf.size(), f)) > 0)
      text.insert(text.end(), buf.begin(), buf.begin() + (long)len);
  }
  std::fclose(f);
  text.push_back('\n');
  std::vector<int64_t> es, ed;  // edges source -> destination
  size_t pos = 0;
  const size_t end = text.size();
  auto line_end = [&](size_t q) {
    while (q < end && text[q] != '\n') q++;
    return q;
  };
  auto parse_int = [&](size_t& q, size_t le, int64_t& v) {
    while (q < le && (text[q] == ' ' || text[q] == '\t' || text[q] == '\r' || text[q] == ',')) q++;
    if (q >= le || text[q] < '0' || text[q] > '9') return false;
    v = 0;
    while (q < le && text[q] >= '0' && text[q] <= '9') v = v * 10 + (text[q++] - '0');
    return true;
  };
  for (int s = 0; s < skip && pos < end; s++) pos = line_end(pos) + 1;
  bool size_line = mtx;  // Matrix Market: the first non-comment line is "rows cols nnz"
  while (pos < end) {
    const size_t le = line_end(pos);
    size_t q = pos;
    while (q < le && (text[q] == ' ' || text[q] == '\t' || text[q] == '\r')) q++;
    const bool blank = q >= le;
    const bool comment = mtx && q < le && text[q] == '%';
    if (!blank && !comment) {
      if (size_line) {
        size_line = false;
      } else {
        int64_t a, b;
        if (!parse_int(q, le, a) || !parse_int(q, le, b)) return TGP_EINVAL;
        if (mtx) {
          if (a < 1 || b < 1) return TGP_EINVAL;
          es.push_back(b - 1);  // column -> row
          ed.push_back(a - 1);
        } else {
          es.push_back(a);
          ed.push_back(b);
        }
      }
    }
    pos = le + 1;
  }
  std::vector<char>().swap(text);
  // vertices that keep an edge once self loops are gone, in id order
  std::vector<int64_t> ids;
  ids.reserve(2 * es.size());
  for (size_t i = 0; i < es.size(); i++)
    if (es[i] != ed[i]) {
      ids.push_back(es[i]);
      ids.push_back(ed[i]);
    }
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  if (ids.empty() || (int64_t)ids.size() > INT32_MAX) return TGP_EINVAL;
  auto rank = [&](int64_t x) {
    return (int32_t)(std::lower_bound(ids.begin(), ids.end(), x) - ids.begin());
  };
  std::vector<int32_t> src, dst;
  src.reserve(2 * es.size());
  dst.reserve(2 * es.size());
  for (size_t i = 0; i < es.size(); i++) {
    if (es[i] == ed[i]) continue;
    const int32_t u = rank(es[i]), v = rank(ed[i]);
    src.push_back(u);
    dst.push_back(v);
    src.push_back(v);
    dst.push_back(u);
  }
  std::vector<int64_t>().swap(es);
  std::vector<int64_t>().swap(ed);
  return build_csr((int64_t)ids.size(), src, dst, 0, 0, out);
}

// The files conversion.py writes (71-110): row.txt, column.txt, weight.txt
// (all ones), output_output.csr and temp.graph (METIS text, edges counted
// once, unit weights), into directory `dir`.
int tgp_write_conversion_files(const tgp_graph* g, const char* dir) {
  if (!g || !dir || g->n <= 0) return TGP_EINVAL;
  const std::string d(dir);
  const int64_t n = g->n, m = g->m;
  auto emit = [&](const std::string& name, auto body) {
    FILE* f = std::fopen((d + "/" + name).c_str(), "wb");
    if (!f) return false;
    std::string buf;
    auto flush = [&](bool force) {
      if (force || buf.size() > (1 << 22)) {
        std::fwrite(buf.data(), 1, buf.size(), f);
        buf.clear();
      }
    };
    body(buf, flush);
    flush(true);
    return std::fclose(f) == 0;
  };
  char num[32];
  auto put = [&](std::string& buf, long long v, const char* sep) {
    const int len = std::snprintf(num, sizeof(num), "%lld%s", v, sep);
    buf.append(num, (size_t)len);
  };
  bool ok = emit("row.txt", [&](std::string& b, auto& fl) {
    for (int64_t i = 0; i <= n; i++) {
      put(b, g->xadj[i], "\n");
      fl(false);
    }
  });
  ok = ok && emit("column.txt", [&](std::string& b, auto& fl) {
    for (int64_t i = 0; i < m; i++) {
      put(b, g->adjncy[i], "\n");
      fl(false);
    }
  });
  ok = ok && emit("weight.txt", [&](std::string& b, auto& fl) {
    for (int64_t i = 0; i < m; i++) {
      b += "1\n";
      fl(false);
    }
  });
  ok = ok && emit("output_output.csr", [&](std::string& b, auto& fl) {
    put(b, n, " ");
    put(b, m, " ");
    put(b, n + 1, " ");
    put(b, m, "\n");
    for (int64_t i = 0; i <= n; i++) {
      put(b, g->xadj[i], " ");
      fl(false);
    }
    b += "\n";
    for (int64_t i = 0; i < m; i++) {
      put(b, g->adjncy[i], " ");
      fl(false);
    }
  });
  ok = ok && emit("temp.graph", [&](std::string& b, auto& fl) {
    put(b, n, " ");
    put(b, m / 2, " ");
    b += "1\n";
    for (int64_t i = 0; i < n; i++) {
      for (int64_t j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
        put(b, (long long)g->adjncy[j] + 1, " 1");
        if (j + 1 < g->xadj[i + 1]) b += ' ';
      }
      b += '\n';
      fl(false);
    }
  });
  return ok ? TGP_OK : TGP_EIO;
}

void tgp_graph_free(tgp_graph* g) {
  if (!g) return;
  std::free((void*)g->xadj);
  std::free((void*)g->adjncy);
  std::free((void*)g->adjwgt);
  g->xadj = nullptr;
  g->adjncy = nullptr;
  g->adjwgt = nullptr;
  g->n = g->m = 0;
}

}  // extern "C"
