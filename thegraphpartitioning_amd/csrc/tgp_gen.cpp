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
#include <cstring>
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
