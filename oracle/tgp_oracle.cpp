// tgp_oracle.cpp -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the multilevel
// k-way partitioning path of iamrupamroyy/TheGraphPartitioning.
//
// A sequential restatement of src/gpu_part.cu under the SEQ schedule: each
// kernel launch behaves as if its threads ran to completion one at a time in
// ascending global thread index, host code as written.  Under SEQ every
// atomic of the reference becomes a deterministic rule:
//   matchVertices CAS (gpu_part.cu:70-77)     -> greedy matching in vertex order
//   rematchVertices (gpu_part.cu:81-106)      -> no-op (all matches mutual)
//   relabel atomicAdd (gpu_part.cu:125)       -> coarse ids in ascending root order
//   bufferCandidateMoves atomicAdd (:343)     -> moves in ascending vertex order
//   selectMoves atomicAdd (:378)              -> rank in post-sort order per destination
// Reference bugs are reproduced on purpose (coarse row r built from fine vertex
// r, mod-64 bucket dedup, 5%+slack balance, last-vertex row end = nedges
// parameter, part counts not restored with the best partition, ...).
//
// Floating point: every fp32 expression is evaluated in the reference's order
// with no contraction (build with -ffp-contract=off, no fast-math).
//
// Loaded only by tests/, the smoke check and bench.py's CPU-baseline leg.

#include "tgp_oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Planted-mistake switch for the tests that show the known-answer tests can
// fail (tgpo_plant_mistake); 0 everywhere else.
int g_plant = 0;

struct Graph {
  int64_t n = 0;
  std::vector<int64_t> xadj;  // n+1
  std::vector<int32_t> adj;   // xadj[n]
  std::vector<int32_t> w;     // xadj[n]
};

struct View {
  int64_t n;
  const int64_t* xadj;
  const int32_t* adj;
  const int32_t* w;
};

View view_of(const Graph& g) { return View{g.n, g.xadj.data(), g.adj.data(), g.w.data()}; }

struct Move {  // gpu_part.cu:32-37
  int32_t v, src, dst;
  float gain;
};

struct Tracer {
  tgpo_trace* buf;
  int64_t cap;
  int64_t len;
  void push(int32_t kind, int32_t level, int32_t phase, int32_t pass, int64_t a,
            int64_t b, int64_t c, int64_t d, uint64_t h) {
    if (buf && len < cap) {
      tgpo_trace& t = buf[len];
      t.kind = kind; t.level = level; t.phase = phase; t.pass = pass;
      t.a = a; t.b = b; t.c = c; t.d = d; t.hash = h;
    }
    len++;
  }
};

uint64_t fnv(const int32_t* v, int64_t n) {
  uint64_t h = 1469598103934665603ull;
  for (int64_t i = 0; i < n; i++) {
    uint32_t x = (uint32_t)v[i];
    for (int b = 0; b < 4; b++) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

// Row end used by the refinement kernels and the edge cut: the last vertex's
// range ends at the caller's `nedges` (gpu_part.cu:277, 324, 408).
inline int64_t row_end_param(const View& g, int64_t v, int64_t nedges) {
  return (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
}

// gpu_part.cu:55-78 with SEQ; rematch (81-106, 1029-1046) is a no-op after it.
void match_seq(const View& g, int32_t* match) {
  const int64_t n = g.n;
  for (int64_t t = 0; t < n; t++) match[t] = -1;            // resetMatch 245-250
  for (int64_t t = 0; t < n; t++) {
    if (match[t] != -1) continue;                          // 58
    int32_t best = -1;
    int32_t maxw = -1;                                     // 62: strict '>' from -1
    for (int64_t i = g.xadj[t]; i < g.xadj[t + 1]; i++) {
      int32_t u = g.adj[i];
      int32_t wt = g.w[i];
      if (u < n && wt > maxw && match[u] == -1 && u != t) {  // 65
        maxw = wt;
        best = u;
      }
    }
    if (best != -1) {                                      // 70-77: CAS succeeds under SEQ
      match[best] = (int32_t)t;
      match[t] = best;
    }
  }
}

// gpu_part.cu:109-118 (mapVertices) + 122-142 (relabel) under SEQ.
int64_t relabel_seq(int64_t n, const int32_t* match, int32_t* new_ids, int32_t* tb) {
  int64_t cnt = 0;
  for (int64_t t = 0; t < n; t++) new_ids[t] = -1;
  for (int64_t t = 0; t < n; t++) {
    const int32_t m = match[t];
    const bool root = (m == -1) || (t < m);                // cmap[t] == t
    if (!root) continue;
    const int64_t id = cnt++;
    new_ids[t] = (int32_t)id;
    tb[2 * id] = (int32_t)t;
    // partner scan 130-135: first i > t with cmap[i] == t, i.e. the mate.
    if (m != -1 && match[m] == t) {
      tb[2 * id + 1] = m;
      new_ids[m] = (int32_t)id;
    } else {
      tb[2 * id + 1] = -1;
    }
  }
  return cnt;
}

// gpu_part.cu:145-242: countUniqueEdges + prefix_sum + fillEdgeArrays +
// updateRowPtr.  Coarse row r is built from FINE vertex r and then fine
// match[r] (only when match[r] < n_c); entries x == new_ids[r] are skipped;
// bucket x%64 is owned by its first entry, later equal x add, others drop.
int64_t contract_seq(const View& g, const int32_t* match, const int32_t* new_ids,
                     int64_t n_c, int64_t* ox, int32_t* oa, int32_t* ow) {
  int64_t pos = 0;
  ox[0] = 0;
  for (int64_t r = 0; r < n_c; r++) {
    const int32_t me = new_ids[r];
    uint64_t mask = 0;
    int32_t owner[64];
    int64_t slot[64];
    auto add = [&](int32_t x, int32_t wt) {
      const int b = x % 64;
      if (!((mask >> b) & 1ull)) {
        mask |= 1ull << b;
        owner[b] = x;
        slot[b] = pos;
        oa[pos] = x;
        ow[pos] = wt;
        pos++;
      } else if (owner[b] == x) {
        ow[slot[b]] = (int32_t)((uint32_t)ow[slot[b]] + (uint32_t)wt);
      }
    };
    for (int64_t i = g.xadj[r]; i < g.xadj[r + 1]; i++) {
      const int32_t x = new_ids[g.adj[i]];
      if (x != me && x != -1) add(x, g.w[i]);
    }
    const int32_t mr = match[r];
    if (mr != -1 && mr < n_c) {
      for (int64_t i = g.xadj[mr]; i < g.xadj[mr + 1]; i++) {
        const int32_t x = new_ids[g.adj[i]];
        if (x != me && x != -1) add(x, g.w[i]);
      }
    }
    ox[r + 1] = pos;
  }
  return pos;
}

// Repaired contraction (no counterpart in the reference; opt-in): coarse row
// c is built from its own fine vertices, the root tb[2c] and then the mate
// tb[2c+1]: their entries in row order, relabelled through new_ids, self
// loops (x == c) removed, each distinct x kept at its first occurrence with
// the weights of all its occurrences summed (int32 wrap).
int64_t contract_repaired(const View& g, const int32_t* tb, const int32_t* new_ids,
                          int64_t n_c, int64_t* ox, int32_t* oa, int32_t* ow) {
  int64_t pos = 0;
  ox[0] = 0;
  std::vector<int64_t> slot_of((size_t)n_c, -1);
  std::vector<int32_t> touched;
  for (int64_t c = 0; c < n_c; c++) {
    touched.clear();
    for (int s = 0; s < 2; s++) {
      const int32_t f = tb[2 * c + s];
      if (f < 0) continue;
      for (int64_t i = g.xadj[f]; i < g.xadj[f + 1]; i++) {
        const int32_t x = new_ids[g.adj[i]];
        if (x == c || x == -1) continue;
        if (slot_of[x] < 0) {
          slot_of[x] = pos;
          touched.push_back(x);
          oa[pos] = x;
          ow[pos] = g.w[i];
          pos++;
        } else {
          if (g_plant != 3)  // planted mistake 3: keep the first weight only
            ow[slot_of[x]] = (int32_t)((uint32_t)ow[slot_of[x]] + (uint32_t)g.w[i]);
        }
      }
    }
    for (int32_t x : touched) slot_of[x] = -1;
    ox[c + 1] = pos;
  }
  return pos;
}

// gpu_part.cu:803-853: the METIS text.  A pair (min,max) comes from every
// arc with 0 <= u < n and u != i using the true row end; its weight is the
// sum over both directions; each line lists neighbours ascending.
std::string metis_text_seq(const View& g, int64_t* npairs) {
  const int64_t n = g.n;
  const int64_t alen = g.xadj[n];
  struct P { int32_t u, v; int32_t w; };
  std::vector<P> arcs;
  for (int64_t i = 0; i < n; i++) {
    const int64_t s = g.xadj[i];
    const int64_t e = (i + 1 < n) ? g.xadj[i + 1] : alen;
    for (int64_t j = s; j < e; j++) {
      const int32_t nb = g.adj[j];
      if (nb >= 0 && nb < n && i != nb) {
        arcs.push_back(P{std::min((int32_t)i, nb), std::max((int32_t)i, nb), g.w[j]});
      }
    }
  }
  std::sort(arcs.begin(), arcs.end(), [](const P& a, const P& b) {
    return a.u != b.u ? a.u < b.u : a.v < b.v;
  });
  std::vector<P> pairs;
  for (const P& a : arcs) {
    if (!pairs.empty() && pairs.back().u == a.u && pairs.back().v == a.v)
      pairs.back().w = (int32_t)((uint32_t)pairs.back().w + (uint32_t)a.w);
    else
      pairs.push_back(a);
  }
  *npairs = (int64_t)pairs.size();
  std::vector<std::vector<std::pair<int32_t, int32_t>>> lines(n);
  for (const P& p : pairs) {
    lines[p.u].push_back({p.v, p.w});
    lines[p.v].push_back({p.u, p.w});
  }
  std::string out = std::to_string(n) + " " + std::to_string(pairs.size()) + " 001\n";
  for (int64_t i = 0; i < n; i++) {
    bool first = true;
    for (const auto& e : lines[i]) {
      if (!first) out += ' ';
      first = false;
      out += std::to_string((int64_t)e.first + 1);
      out += ' ';
      out += std::to_string(e.second);
    }
    out += '\n';
  }
  return out;
}

// gpu_part.cu:400-422 + 468: int32 sum with wrap, then /2 (truncating).
int64_t cut_seq(const View& g, int64_t nedges, const int32_t* part, int32_t k, int wide) {
  const int64_t n = g.n;
  uint64_t s = 0;
#pragma omp parallel for reduction(+ : s) schedule(dynamic, 1024)
  for (int64_t v = 0; v < n; v++) {
    const int64_t st = g.xadj[v];
    const int64_t en = row_end_param(g, v, nedges);
    const int32_t pv = part[v];
    for (int64_t i = st; i < en; i++) {
      const int32_t u = g.adj[i];
      if (u < n) {
        const int32_t pu = part[u];
        if (pv != pu && pv >= 0 && pv < k && pu >= 0 && pu < k) s += (uint64_t)(int64_t)g.w[i];
      }
    }
  }
  if (!wide) {
    const int32_t c = (int32_t)(uint32_t)s;
    return (int64_t)(c / 2);
  }
  return (int64_t)s / 2;
}

// gpu_part.cu:258-305 (gains, both k<=32 and k>32 paths have the same
// semantics: fp32 tallies accumulated in row order).
void gains_seq(const View& g, int64_t nedges, const int32_t* part, const int32_t* pc,
               int32_t k, int phase, float* gain, int32_t* target,
               std::vector<float>& counts) {
  const int64_t n = g.n;
  const int64_t T = n / k;
  const float Tf = (float)T;
  (void)counts;
  // vertices are independent (each writes only its own gain and target)
#pragma omp parallel
  {
  std::vector<float> cnt((size_t)k, 0.0f);
#pragma omp for schedule(dynamic, 256)
  for (int64_t v = 0; v < n; v++) {
    const int32_t cur = part[v];
    if (cur >= k || cur < 0) continue;                      // 262
    std::fill(cnt.begin(), cnt.end(), 0.0f);
    const int64_t st = g.xadj[v];
    const int64_t en = row_end_param(g, v, nedges);
    for (int64_t i = st; i < en; i++) {
      const int32_t u = g.adj[i];
      if (u < n) {
        const int32_t p = part[u];
        if (p < k && p >= 0) cnt[p] += (float)g.w[i];
      }
    }
    float best = -1e9f;
    int32_t tgt = cur;
    const float srcSize = (float)pc[cur];
    for (int32_t q = 0; q < k; q++) {
      if (q == cur) continue;
      const float edgeGain = cnt[q] - cnt[cur];
      const float destSize = (float)pc[q];
      float bg = phase ? 500.0f * (Tf - destSize) / Tf : 200.0f * (Tf - destSize) / Tf;
      if (srcSize > Tf * 1.05f) bg += phase ? 250.0f : 100.0f;
      if (destSize < Tf * 0.95f) bg += phase ? 125.0f : 50.0f;
      const float gg = edgeGain + bg;
      if (gg > best || (gg == best && q < tgt)) {
        best = gg;
        tgt = q;
      }
    }
    gain[v] = best;
    target[v] = tgt;
  }
  }
}

// gpu_part.cu:472-484 executed for the stage sequence of 487-506.
template <class E, class Key>
void bitonic_seq(E* d, int64_t len, Key key) {
  for (int64_t kk = 2; kk <= len; kk <<= 1) {
    for (int64_t j = kk >> 1; j > 0; j >>= 1) {
      // the pairs of one stage are disjoint
#pragma omp parallel for schedule(static) if (len >= 65536)
      for (int64_t idx = 0; idx < len; idx++) {
        const int64_t ixj = idx ^ j;
        if (ixj > idx || ixj >= len) continue;
        const bool up = ((idx & kk) == 0);
        const bool swp = (key(d[idx]) < key(d[ixj])) ^ up;
        if (swp) std::swap(d[idx], d[ixj]);
      }
    }
  }
}

// gpu_part.cu:509-759.
void refine_seq(const View& g, int64_t nedges, int32_t k, int wide, int32_t* part,
                int32_t* pc, int64_t* edgeCut, Tracer& tr, int level) {
  const int64_t n = g.n;
  const int64_t maxMoves = n;
  std::vector<float> gain((size_t)n, 0.0f);
  std::vector<int32_t> target((size_t)n, 0);
  std::vector<uint8_t> cand((size_t)n, 0), fin((size_t)n, 0);
  std::vector<Move> buf((size_t)maxMoves, Move{0, 0, 0, 0.0f});
  std::vector<uint8_t> sel((size_t)maxMoves, 0);
  std::vector<int64_t> outc((size_t)k), inc((size_t)k), avail((size_t)k), dc((size_t)k);
  std::vector<float> counts;

  int64_t bestCut = *edgeCut;
  float bestScore = 1e9f;
  std::vector<int32_t> best(part, part + n);
  const int64_t T = n / k;
  const float Tf = (float)T;
  const float maxImbalance = 0.05f;
  int total_passes = 0;

  for (int phase = 0; phase < 2; phase++) {
    const int maxPass = phase ? 40 : 30;
    bool balanced = false;
    for (int pass = 0; pass < maxPass && !(phase && balanced); pass++) {
      std::fill(outc.begin(), outc.end(), 0);
      std::fill(inc.begin(), inc.end(), 0);
      gains_seq(g, nedges, part, pc, k, phase, gain.data(), target.data(), counts);
#pragma omp parallel for schedule(static)
      for (int64_t v = 0; v < n; v++)                       // markCandidates 308-312
        cand[v] = phase ? (gain[v] > -50.0f) : (gain[v] > -10.0f);
#pragma omp parallel for schedule(dynamic, 1024)
      for (int64_t v = 0; v < n; v++) {                     // tokenPassing 316-333
        if (!cand[v]) { fin[v] = 0; continue; }
        bool win = true;
        const int64_t st = g.xadj[v];
        const int64_t en = row_end_param(g, v, nedges);
        for (int64_t i = st; i < en; i++) {
          const int32_t u = g.adj[i];
          if (u < n && cand[u] &&
              ((gain[u] > gain[v] + 10.0f) || (gain[u] == gain[v] && u < v))) {
            win = false;
            break;
          }
        }
        fin[v] = win;
      }
      int64_t M = 0;                                        // bufferCandidateMoves 337-350
      for (int64_t v = 0; v < n; v++) {
        if (!fin[v]) continue;
        const int32_t src = part[v], dst = target[v];
        if (src < k && dst < k && src != dst && src >= 0 && dst >= 0) {
          if (M < maxMoves) {
            buf[M] = Move{(int32_t)v, src, dst, gain[v]};
            outc[src]++;
            inc[dst]++;
          }
          M++;
        }
      }
      M = std::min(M, maxMoves);
      total_passes++;
      if (M <= 0) break;                                    // 638-640 / 742-744

      for (int32_t q = 0; q < k; q++) {                     // computeAvailableCapacity 357-369
        const int64_t current = pc[q];
        const int64_t slack = std::max<int64_t>(100, n / ((int64_t)k * 10));
        const int64_t imb = (int64_t)(Tf * maxImbalance);
        const int64_t maxSize = T + imb + slack;
        const int64_t minSize = T - imb - slack;
        int64_t a = T - current + outc[q] + slack;
        if (a < 0) a = 0;
        if (a > maxSize - current + slack) a = maxSize - current + slack;
        if (current <= minSize && !phase)  // planted mistake 1: slack, not slack/2
          a = std::min(a, g_plant == 1 ? slack : slack / 2);
        avail[q] = a;
      }

      int64_t len = M;                                      // customSortMoves 487-506
      int64_t padded = 1;
      while (padded < len) padded <<= 1;
      if (padded > maxMoves) padded = maxMoves;
      if (padded > len) {
        for (int64_t i = len; i < padded; i++) buf[i] = Move{-1, -1, -1, -1e9f};
        len = padded;
      }
      if (g_plant != 2)  // planted mistake 2: rank moves in buffer order (no sort)
        bitonic_seq(buf.data(), len, [](const Move& m) { return m.gain; });

      std::fill(dc.begin(), dc.end(), 0);                   // selectMoves 373-380
      int64_t pads = 0, S = 0;
      for (int64_t i = 0; i < M; i++) {
        const int32_t dst = buf[i].dst;
        if (dst < 0 || dst >= k) {                          // padding in the prefix: not selected
          sel[i] = 0;
          pads++;
          continue;
        }
        const int64_t pos = dc[dst]++;
        sel[i] = pos < avail[dst];
      }
      for (int64_t i = 0; i < M; i++) {                     // commitMoves 383-397
        if (!sel[i]) continue;
        S++;
        const Move m = buf[i];
        if (m.src < k && m.dst < k && m.src >= 0 && m.dst >= 0) {
          const int32_t old = part[m.v];
          part[m.v] = m.dst;
          if (old == m.src) {
            pc[m.dst]++;
            pc[m.src]--;
          } else {
            part[m.v] = old;
          }
        }
      }
      const int64_t cur = cut_seq(g, nedges, part, k, wide);
      float score = 0.0f;                                   // 621-626 / 723-731
      balanced = true;
      for (int32_t q = 0; q < k; q++) {
        const float diff = (float)pc[q] - Tf;
        const float sq = diff * diff;
        score = score + sq;
        if (std::fabs(diff) > Tf * maxImbalance) balanced = false;
      }
      score = std::sqrt(score / (float)k);
      if (cur < bestCut || ((float)cur <= (float)bestCut * 1.01f && score < bestScore)) {
        bestCut = cur;
        bestScore = score;
        std::copy(part, part + n, best.begin());
      }
      *edgeCut = cur;
      tr.push(TGPO_TR_PASS, level, phase, pass, M, S, cur, pads, fnv(part, n));
      const bool quiet = (M < n / 200) && (S < n / 400);   // 634 / 738
      if (quiet && (phase == 0 || balanced)) break;
    }
    std::copy(best.begin(), best.end(), part);              // 642 / 746 (part counts untouched)
    if (phase == 0) *edgeCut = bestCut;
  }
  *edgeCut = cut_seq(g, nedges, part, k, wide);             // 749
  tr.push(TGPO_TR_REFINE, level, 0, total_passes, n, nedges, *edgeCut, 0, fnv(part, n));
}

bool valid_csr(int64_t n, const int64_t* xadj, const int32_t* adj) {
  if (n <= 0 || xadj[0] != 0) return false;
  for (int64_t i = 0; i < n; i++)
    if (xadj[i + 1] < xadj[i]) return false;
  if (xadj[n] <= 0) return false;
  for (int64_t i = 0; i < xadj[n]; i++)
    if (adj[i] < 0 || adj[i] >= n) return false;
  return true;
}

// Reads the part file the way gpu_part.cu:1163-1174 does: ifstream >> int,
// a failed read leaves 0, out-of-range values become 0.
int read_part_file(const std::string& path, int64_t n_c, int32_t k, std::vector<int32_t>& p) {
  std::ifstream fin(path);
  if (!fin.is_open()) return TGPO_EPARTFILE;
  std::vector<int> h((size_t)n_c, 0);
  for (int64_t i = 0; i < n_c; i++) {
    fin >> h[i];
    if (h[i] < 0 || h[i] >= k) h[i] = 0;
  }
  p.assign(h.begin(), h.end());
  return TGPO_OK;
}

}  // namespace

extern "C" {

void tgpo_plant_mistake(int32_t which) { g_plant = which; }

int32_t tgpo_set_threads(int32_t threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  return omp_get_max_threads();
#else
  (void)threads;
  return 1;
#endif
}

uint64_t tgpo_hash(const int32_t* v, int64_t n) { return fnv(v, n); }

int tgpo_match(int64_t n, const int64_t* xadj, const int32_t* adjncy,
               const int32_t* adjwgt, int32_t* match) {
  if (n <= 0) return TGPO_EINVAL;
  match_seq(View{n, xadj, adjncy, adjwgt}, match);
  return TGPO_OK;
}

int64_t tgpo_relabel(int64_t n, const int32_t* match, int32_t* new_ids, int32_t* tb) {
  return relabel_seq(n, match, new_ids, tb);
}

int64_t tgpo_contract(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                      const int32_t* adjwgt, const int32_t* match,
                      const int32_t* new_ids, int64_t n_c, int64_t* ox,
                      int32_t* oa, int32_t* ow) {
  return contract_seq(View{n, xadj, adjncy, adjwgt}, match, new_ids, n_c, ox, oa, ow);
}

int64_t tgpo_contract_repaired(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                               const int32_t* adjwgt, const int32_t* tb,
                               const int32_t* new_ids, int64_t n_c, int64_t* ox,
                               int32_t* oa, int32_t* ow) {
  return contract_repaired(View{n, xadj, adjncy, adjwgt}, tb, new_ids, n_c, ox, oa, ow);
}

int64_t tgpo_metis_text(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                        const int32_t* adjwgt, char* buf, int64_t cap, int64_t* npairs) {
  std::string s = metis_text_seq(View{n, xadj, adjncy, adjwgt}, npairs);
  if (buf && cap > 0) std::memcpy(buf, s.data(), (size_t)std::min<int64_t>(cap, (int64_t)s.size()));
  return (int64_t)s.size();
}

int64_t tgpo_edge_cut(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                      const int32_t* adjwgt, int64_t nedges, const int32_t* part,
                      int32_t k, int32_t wide) {
  return cut_seq(View{n, xadj, adjncy, adjwgt}, nedges, part, k, wide);
}

int tgpo_gains(int64_t n, const int64_t* xadj, const int32_t* adjncy,
               const int32_t* adjwgt, int64_t nedges, const int32_t* part,
               const int32_t* pc, int32_t k, int32_t phase, float* gain,
               int32_t* target) {
  if (n <= 0 || k <= 0) return TGPO_EINVAL;
  std::vector<float> counts;
  gains_seq(View{n, xadj, adjncy, adjwgt}, nedges, part, pc, k, phase, gain, target, counts);
  return TGPO_OK;
}

int tgpo_bitonic(float* gain, int32_t* id, int64_t len) {
  struct E { float g; int32_t i; };
  std::vector<E> e((size_t)len);
  for (int64_t i = 0; i < len; i++) e[i] = E{gain[i], id[i]};
  bitonic_seq(e.data(), len, [](const E& x) { return x.g; });
  for (int64_t i = 0; i < len; i++) { gain[i] = e[i].g; id[i] = e[i].i; }
  return TGPO_OK;
}

int tgpo_refine(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                const int32_t* adjwgt, int64_t nedges, int32_t k, int32_t wide,
                int32_t* part, int32_t* pc, int64_t* cut, tgpo_trace* trace,
                int64_t trace_cap, int64_t* trace_len) {
  if (n <= 0 || k <= 0) return TGPO_EINVAL;
  Tracer tr{trace, trace_cap, 0};
  refine_seq(View{n, xadj, adjncy, adjwgt}, nedges, k, wide, part, pc, cut, tr, 0);
  if (trace_len) *trace_len = tr.len;
  return TGPO_OK;
}

int64_t tgpo_final_text(int64_t n, const int32_t* part, int32_t k, char* buf, int64_t cap) {
  std::vector<std::vector<int64_t>> by((size_t)k);
  for (int64_t i = 0; i < n; i++) {
    const int32_t p = part[i];
    if (p >= 0 && p < k) by[p].push_back(i);
    else by[0].push_back(i);
  }
  std::string s;
  for (int32_t q = 0; q < k; q++) {
    s += "Partition " + std::to_string(q) + ": ";
    for (size_t i = 0; i < by[q].size(); i++) {
      s += std::to_string(by[q][i]);
      if (i + 1 < by[q].size()) s += ',';
    }
    s += '\n';
  }
  if (buf && cap > 0) std::memcpy(buf, s.data(), (size_t)std::min<int64_t>(cap, (int64_t)s.size()));
  return (int64_t)s.size();
}

int tgpo_partition(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                   const int32_t* adjwgt, const tgpo_options* opt, int32_t* part,
                   int32_t* part_count, int64_t* edge_cut, tgpo_trace* trace,
                   int64_t trace_cap, int64_t* trace_len) {
  if (!opt || opt->k <= 0 || n <= 0 || !xadj || !adjncy || !adjwgt) return TGPO_EINVAL;
  if (!valid_csr(n, xadj, adjncy)) return TGPO_EINVAL;
  const int32_t k = opt->k;
  const int wide = opt->wide;
  const int64_t m = xadj[n];
  if (!wide) {                                              // B11
    if (n > INT32_MAX || m > INT32_MAX) return TGPO_ELIMIT;
    if (n * (int64_t)k > (int64_t)INT32_MAX) return TGPO_ELIMIT;  // gpu_part.cu:531
  }
  Tracer tr{trace, trace_cap, 0};

  std::vector<Graph> L(1);
  L[0].n = n;
  L[0].xadj.assign(xadj, xadj + n + 1);
  L[0].adj.assign(adjncy, adjncy + m);
  L[0].w.assign(adjwgt, adjwgt + m);
  std::vector<int64_t> ecount{m};
  std::vector<std::vector<int32_t>> tbh(1);

  const int64_t target = opt->coarsen_floor > 0 ? opt->coarsen_floor
                                                : std::max<int64_t>(500, n / 500);  // 897
  int64_t ncur = n;
  std::vector<int32_t> match, new_ids, tb;
  while (ncur > target) {                                   // 1003-1152
    // repaired mode: a contraction that kept more than 93 % of the vertices
    // ends the coarsening before the next matching (whose level would keep
    // more than 95 % on power-law graphs and be dropped by the rule below)
    if (opt->repair && L.size() >= 2 && 100 * L.back().n > 93 * L[L.size() - 2].n) break;
    const View g = view_of(L.back());
    match.resize(ncur);
    new_ids.resize(ncur);
    tb.assign(2 * ncur, -1);
    match_seq(g, match.data());
    const int64_t n_c = relabel_seq(ncur, match.data(), new_ids.data(), tb.data());
    if (n_c <= 0 || n_c >= ncur) break;                     // 1069
    if (opt->repair && 20 * n_c > 19 * ncur) break;         // repaired mode: stall stop
    Graph c;
    c.n = n_c;
    c.xadj.resize(n_c + 1);
    const int64_t cap = 2 * g.xadj[ncur];
    c.adj.resize(cap);
    c.w.resize(cap);
    const int64_t m_c =
        opt->repair ? contract_repaired(g, tb.data(), new_ids.data(), n_c, c.xadj.data(),
                                        c.adj.data(), c.w.data())
                    : contract_seq(g, match.data(), new_ids.data(), n_c, c.xadj.data(),
                                   c.adj.data(), c.w.data());
    int64_t max_edges = 2 * ecount.back();
    if (!wide && max_edges > INT32_MAX) max_edges = INT32_MAX;
    tr.push(TGPO_TR_LEVEL, (int32_t)L.size() - 1, 0, 0, ncur, g.xadj[ncur], n_c, m_c, 0);
    if (m_c <= 0 || m_c > max_edges) {                      // 1121-1125
      if (trace_len) *trace_len = tr.len;
      return TGPO_EEDGES;
    }
    c.adj.resize(m_c);
    c.w.resize(m_c);
    tb.resize(2 * n_c);
    L.push_back(std::move(c));
    ecount.push_back(m_c);
    tbh.push_back(tb);
    ncur = n_c;
  }
  const int pass = (int)L.size() - 1;

  // Initial partition on the coarsest graph (1157-1185).
  const View cg = view_of(L[pass]);
  int64_t npairs = 0;
  const std::string text = metis_text_seq(cg, &npairs);
  std::vector<int32_t> pcur;
  const int64_t n_c = cg.n;
  if (opt->init_mode == TGPO_INIT_BLOCK) {
    pcur.resize(n_c);
    for (int64_t i = 0; i < n_c; i++) pcur[i] = (int32_t)((i * (int64_t)k) / n_c);
  } else if (opt->init_mode == TGPO_INIT_ARRAY) {
    pcur.assign((size_t)n_c, 0);
    for (int64_t i = 0; i < n_c && i < opt->init_part_len; i++) pcur[i] = opt->init_part[i];
  } else if (opt->init_mode == TGPO_INIT_GPMETIS) {
    {
      std::ofstream out("output.graph");
      if (!out.is_open()) return TGPO_EINITPART;
      out << text;
    }
    const std::string cmd = "gpmetis -ptype=rb ./output.graph " + std::to_string(k);
    const int ret = std::system(cmd.c_str());
    if (ret != 0) return TGPO_EINITPART;
    const int rc = read_part_file("output.graph.part." + std::to_string(k), n_c, k, pcur);
    if (rc != TGPO_OK) return rc;
  } else {
    return TGPO_EINVAL;
  }
  for (auto& p : pcur)
    if (p < 0 || p >= k) p = 0;
  std::vector<int32_t> pc((size_t)k, 0);
  for (int64_t i = 0; i < n_c; i++) pc[pcur[i]]++;
  int64_t cut = cut_seq(cg, npairs, pcur.data(), k, wide);  // 1191 (B5: nedges = |pairs|)
  tr.push(TGPO_TR_INIT, pass, 0, 0, n_c, npairs, cut, 0, fnv(pcur.data(), n_c));
  refine_seq(cg, npairs, k, wide, pcur.data(), pc.data(), &cut, tr, pass);

  for (int level = pass - 1; level >= 0; level--) {        // 1205-1274
    const View g = view_of(L[level]);
    const std::vector<int32_t>& t = tbh[level + 1];
    const int64_t ncoarse = L[level + 1].n;
    std::vector<int32_t> pf((size_t)g.n, -1);
    for (int64_t c = 0; c < ncoarse; c++) {                 // kwayMapCoarseToFine 763-772
      const int32_t cp = pcur[c];
      if (cp >= k || cp < 0) continue;
      const int32_t v1 = t[2 * c], v2 = t[2 * c + 1];
      if (v1 >= 0 && v1 < g.n) pf[v1] = cp;
      if (v2 >= 0 && v2 < g.n) pf[v2] = cp;
    }
    std::fill(pc.begin(), pc.end(), 0);                     // 1244-1252
    for (int64_t i = 0; i < g.n; i++) {
      if (pf[i] >= 0 && pf[i] < k) pc[pf[i]]++;
      else { pf[i] = 0; pc[0]++; }
    }
    pcur.swap(pf);
    cut = cut_seq(g, ecount[level], pcur.data(), k, wide);
    refine_seq(g, ecount[level], k, wide, pcur.data(), pc.data(), &cut, tr, level);
  }
  const int64_t fcut = cut_seq(view_of(L[0]), m, pcur.data(), k, wide);  // 1285
  tr.push(TGPO_TR_FINAL, 0, 0, 0, n, m, fcut, pass, fnv(pcur.data(), n));
  std::copy(pcur.begin(), pcur.end(), part);
  std::copy(pc.begin(), pc.end(), part_count);
  *edge_cut = fcut;
  if (trace_len) *trace_len = tr.len;
  return TGPO_OK;
}

}  // extern "C"
