// tgp_balance.hip -- the balance stage of the forms that report at an
// imbalance bound (no counterpart in the reference, whose refinement enforces
// only 5 % + slack on the counts of each level and ends far above it on
// power-law graphs).  At the finest level, with cap = floor((1 + eps) *
// ceil(n / k)):
//
//   balance round: every vertex v of a part p holding more than cap takes the
//     part q != p with room (count < cap) that maximises the exact integer
//     gain T[v][q] - T[v][p] (lowest q on ties); candidates are ranked per
//     destination by (gain desc, v asc) and kept while the rank is below q's
//     room, the kept ones ranked per source the same way and kept while the
//     rank is below p's excess; the kept moves are applied.  Rounds repeat
//     until no part exceeds cap.
//   refine pass: every vertex takes the part q with room that maximises the
//     exact gain; it is a candidate when the gain is positive, and a winner
//     when no neighbouring candidate has a larger gain (or the same gain and a
//     smaller id); winners are ranked per destination by (gain desc, v asc)
//     and applied while the rank is below q's room.  Passes repeat while moves
//     are applied (at most a fixed number).
//
// T[v][q] is the weight of v's arcs (self loops excluded) into part q,
// gathered from the partition by k_bal_tally.  All ranks come from a stable
// radix sort of (part, gain) keys over candidates listed in ascending vertex
// order, so every step is deterministic.
#include "tgp_internal.h"

#include <hipcub/device/device_radix_sort.hpp>

#include <climits>

namespace tgp {
namespace {

constexpr int BAL_BS = 256;
constexpr int BAL_WAVES = BAL_BS / 64;

__device__ __forceinline__ int bl_lane() { return (int)(threadIdx.x & 63); }

// T[v][q] for every v: one wave per vertex, its row tallied in LDS; the
// block's four waves take four consecutive vertices per step.
__global__ __launch_bounds__(BAL_BS) void k_bal_tally(DGraph g, const i32* part, i32 k, i32* T) {
  extern __shared__ i32 sh[];
  const int lane = bl_lane(), wv = threadIdx.x >> 6;
  i32* t = sh + wv * k;
  for (i64 base = (i64)blockIdx.x * BAL_WAVES; base < g.n; base += (i64)gridDim.x * BAL_WAVES) {
    const i64 v = base + wv;
    for (int q = lane; q < k; q += 64) t[q] = 0;
    __syncthreads();
    if (v < g.n) {
      const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
      for (i64 i = s0 + lane; i < s1; i += 64) {
        const i32 u = g.adj[i];
        if (u == (i32)v || u < 0 || u >= g.n) continue;
        const i32 p = part[u];
        if (p >= 0 && p < k) atomicAdd(&t[p], g.w[i]);
      }
    }
    __syncthreads();
    if (v < g.n)
      for (int q = lane; q < k; q += 64) T[v * k + q] = t[q];
    __syncthreads();
  }
}

// Best destination with room for each vertex (mode 0: vertices of parts over
// cap, any gain; mode 1: every vertex, gain > 0).  cand[v] = 1 for a
// candidate, with target[v] and gain[v].  Four lanes per vertex, each taking
// every fourth part of the row; the lanes' bests are combined with the same
// rule (larger gain, then lower part).
__global__ __launch_bounds__(BAL_BS) void k_bal_best(i64 n, const i32* part, const i32* T, i32 k,
                                                     const i32* room, int mode, u8* cand,
                                                     i32* target, i32* gain) {
  const int gl = threadIdx.x & 3;
  const i64 ngroups = (i64)gridDim.x * (BAL_BS / 4);
  for (i64 v = (i64)blockIdx.x * (BAL_BS / 4) + threadIdx.x / 4; v < n; v += ngroups) {
    const i32 p = part[v];
    i32 bq = -1;
    i64 bg = LLONG_MIN;
    const bool look = p >= 0 && p < k && (mode == 1 || room[p] < 0);
    if (look) {
      const i32* Tv = T + v * k;
      const i64 own = Tv[p];
      for (i32 q = gl; q < k; q += 4) {
        if (q == p || room[q] <= 0) continue;
        const i64 gq = (i64)Tv[q] - own;
        if (gq > bg) {
          bg = gq;
          bq = q;
        }
      }
    }
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
      const i64 og = __shfl_xor(bg, o, 4);
      const i32 oq = __shfl_xor(bq, o, 4);
      if (oq >= 0 && (bq < 0 || og > bg || (og == bg && oq < bq))) {
        bg = og;
        bq = oq;
      }
    }
    if (gl == 0) {
      const bool c = look && bq >= 0 && (mode == 0 || bg > 0);
      cand[v] = c;
      target[v] = c ? bq : -1;
      gain[v] = c ? (i32)(bg > INT_MAX ? INT_MAX : bg < INT_MIN ? INT_MIN : bg) : 0;
    }
  }
}

// Tallies after a round: each kept vertex v moving from part[v] to target[v]
// shifts the weight of each of its arcs (v, u) from T[u][part[v]] to
// T[u][target[v]] (the reverse arc (u, v) has the same weight on these
// levels).  Run before the moves are applied.  Each wave takes 64 vertices
// and walks the rows of the kept ones with the whole wave.
__global__ __launch_bounds__(BAL_BS) void k_bal_tally_update(DGraph g, const u8* keep,
                                                             const i32* part,
                                                             const i32* target, i32 k, i32* T) {
  const int lane = bl_lane();
  const i64 gw = ((i64)blockIdx.x * BAL_BS + threadIdx.x) >> 6;
  const i64 nw = ((i64)gridDim.x * BAL_BS) >> 6;
  for (i64 v0 = gw * 64; v0 < g.n; v0 += nw * 64) {
    const i64 mv = v0 + lane;
    u64 todo = __ballot(mv < g.n && keep[mv]);
    while (todo) {
      const int l = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const i64 v = v0 + l;
      const i32 src = part[v], dst = target[v];
      const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
      for (i64 i = s0 + lane; i < s1; i += 64) {
        const i32 u = g.adj[i];
        if (u == (i32)v || u < 0 || u >= g.n) continue;
        const i32 w = g.w[i];
        atomicAdd(&T[(i64)u * k + src], -w);
        atomicAdd(&T[(i64)u * k + dst], w);
      }
    }
  }
}

// Independent winners of the refine pass: no neighbouring candidate with a
// larger gain, or the same gain and a smaller id.  One wave per vertex.
__global__ __launch_bounds__(BAL_BS) void k_bal_token(DGraph g, const u8* cand, const i32* gain,
                                                      u8* win) {
  const int lane = bl_lane(), wv = threadIdx.x >> 6;
  const i64 nw = (i64)gridDim.x * BAL_WAVES;
  for (i64 v = (i64)blockIdx.x * BAL_WAVES + wv; v < g.n; v += nw) {
    if (!cand[v]) {
      if (lane == 0) win[v] = 0;
      continue;
    }
    const i32 gv = gain[v];
    bool lose = false;
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 b = s0; b < s1 && !lose; b += 64) {  // uniform over the wave
      const i64 i = b + lane;
      bool l = false;
      if (i < s1) {
        const i32 u = g.adj[i];
        if (u != (i32)v && u >= 0 && u < g.n && cand[u]) {
          const i32 gu = gain[u];
          l = gu > gv || (gu == gv && u < (i32)v);
        }
      }
      lose = __ballot(l) != 0;
    }
    if (lane == 0) win[v] = !lose;
  }
}

// Candidates with flag set, listed in ascending vertex order: per-block
// counts, then (after an exclusive scan on the host side of the launcher)
// the ordered write of (key, v) pairs with key = part << 32 | (INT_MAX - gain).
__global__ __launch_bounds__(BAL_BS) void k_bal_count(const u8* flag, i64 n, i64 per, i32* bcnt) {
  __shared__ i32 sw[BAL_WAVES];
  const i64 b0 = (i64)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  i32 c = 0;
  for (i64 v = b0 + threadIdx.x; v < b1; v += BAL_BS) c += flag[v] != 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (bl_lane() == 0) sw[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    i32 t = 0;
    for (int w = 0; w < BAL_WAVES; w++) t += sw[w];
    bcnt[blockIdx.x] = t;
  }
}
__global__ void k_bal_scan(i32* bcnt, i32 nb, i64* total) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  i64 run = 0;
  for (i32 b = 0; b < nb; b++) {
    const i32 c = bcnt[b];
    bcnt[b] = (i32)run;
    run += c;
  }
  *total = run;
}
// keypart[v] = the part the key is formed from (target for destination
// ranks, part for source ranks)
__global__ __launch_bounds__(BAL_BS) void k_bal_list(const u8* flag, i64 n, i64 per,
                                                     const i32* boff, const i32* keypart,
                                                     const i32* gain, u64* keys, i32* vals) {
  __shared__ i32 sw[BAL_WAVES];
  const int lane = bl_lane(), wv = threadIdx.x >> 6;
  const i64 b0 = (i64)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  i64 run = boff[blockIdx.x];
  for (i64 base = b0; base < b1; base += BAL_BS) {
    const i64 v = base + threadIdx.x;
    const bool f = v < b1 && flag[v];
    const u64 bm = __ballot(f);
    if (lane == 0) sw[wv] = __popcll(bm);
    __syncthreads();
    i32 before = 0, tot = 0;
    for (int w = 0; w < BAL_WAVES; w++) {
      if (w < wv) before += sw[w];
      tot += sw[w];
    }
    if (f) {
      const u64 below = lane ? (~0ull >> (64 - lane)) : 0ull;
      const i64 o = run + before + __popcll(bm & below);
      keys[o] = ((u64)(u32)keypart[v] << 32) | (u64)(u32)((i64)INT_MAX - (i64)gain[v]);
      vals[o] = (i32)v;
    }
    run += tot;
    __syncthreads();
  }
}
// segment starts of the sorted keys: start[q] = first position with part q
// (filled with cnt on entry, which stays for parts that have none)
__global__ void k_bal_starts(const u64* keys, i64 cnt, i64* start) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (i64)gridDim.x * blockDim.x) {
    const i32 q = (i32)(keys[i] >> 32);
    if (i == 0 || (i32)(keys[i - 1] >> 32) != q) start[q] = i;
  }
}
// keep[v] = rank within its part's segment < limit[part]; cleared for the rest
__global__ void k_bal_keep(const u64* keys, const i32* vals, i64 cnt, const i64* start,
                           const i32* limit, u8* keep) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < cnt;
       i += (i64)gridDim.x * blockDim.x) {
    const i32 q = (i32)(keys[i] >> 32);
    keep[vals[i]] = (i - start[q]) < (i64)limit[q];
  }
}
__global__ void k_bal_apply(const u8* keep, i64 n, const i32* target, i32* part, i32* pc,
                            u64* moved) {
  u64 c = 0;
  for (i64 v = (i64)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (i64)gridDim.x * blockDim.x) {
    if (!keep[v]) continue;
    const i32 p = part[v], q = target[v];
    part[v] = q;
    atomicAdd(&pc[p], -1);
    atomicAdd(&pc[q], 1);
    c++;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (bl_lane() == 0 && c) atomicAdd((unsigned long long*)moved, (unsigned long long)c);
}
__global__ void k_bal_fill_i64(i64* p, i64 n, i64 v) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (i64)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_bal_clear_u8(u8* p, i64 n) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (i64)gridDim.x * blockDim.x)
    p[i] = 0;
}

// avail[q] = max(cap - pc[q], 0): the capacity of the learned passes when the
// form reports at an imbalance bound
__global__ void k_bal_avail(const i32* pc, i32 k, i64 cap, i64* avail) {
  for (i32 q = threadIdx.x; q < k; q += blockDim.x) {
    const i64 a = cap - (i64)pc[q];
    avail[q] = a > 0 ? a : 0;
  }
}

unsigned bl_grid(i64 n, int per, unsigned cap) {
  i64 b = (n + per - 1) / per;
  if (b < 1) b = 1;
  if (b > (i64)cap) b = cap;
  return (unsigned)b;
}

}  // namespace

i64 balance_sort_bytes(i64 max_items) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const u64*)nullptr, (u64*)nullptr,
                                     (const i32*)nullptr, (i32*)nullptr, (int)max_items, 0, 64);
  return (i64)bytes;
}

void launch_bal_tally(DGraph g, const i32* part, i32 k, i32* T, hipStream_t s) {
  k_bal_tally<<<bl_grid(g.n, BAL_WAVES, 1u << 16), BAL_BS, (size_t)BAL_WAVES * k * 4, s>>>(
      g, part, k, T);
}

void launch_bal_best(i64 n, const i32* part, const i32* T, i32 k, const i32* room, int mode,
                     u8* cand, i32* target, i32* gain, hipStream_t s) {
  k_bal_best<<<bl_grid(n, BAL_BS / 4, 1u << 16), BAL_BS, 0, s>>>(n, part, T, k, room, mode,
                                                                 cand, target, gain);
}

void launch_bal_tally_update(DGraph g, const u8* keep, const i32* part, const i32* target, i32 k,
                             i32* T, hipStream_t s) {
  k_bal_tally_update<<<bl_grid(g.n, BAL_BS, 1u << 14), BAL_BS, 0, s>>>(g, keep, part, target, k,
                                                                      T);
}

void launch_bal_token(DGraph g, const u8* cand, const i32* gain, u8* win, hipStream_t s) {
  k_bal_token<<<bl_grid(g.n, BAL_WAVES, 1u << 16), BAL_BS, 0, s>>>(g, cand, gain, win);
}

// keep[v] = 1 for the flagged vertices whose rank by (gain desc, v asc) within
// keypart[v] is below limit[keypart[v]]; 0 elsewhere.  Scratch: BalScratch.
void launch_bal_rank_keep(const u8* flag, i64 n, const i32* keypart, const i32* gain, i32 k,
                          const i32* limit, u8* keep, const BalScratch& S, hipStream_t s) {
  i64 nb = (n + 4095) / 4096;
  if (nb < 1) nb = 1;
  if (nb > 4096) nb = 4096;
  i64 per = (n + nb - 1) / nb;
  per = (per + BAL_BS - 1) / BAL_BS * BAL_BS;
  nb = (n + per - 1) / per;
  if (nb < 1) nb = 1;
  k_bal_count<<<(unsigned)nb, BAL_BS, 0, s>>>(flag, n, per, S.bcnt);
  k_bal_scan<<<1, 64, 0, s>>>(S.bcnt, (i32)nb, S.total);
  k_bal_list<<<(unsigned)nb, BAL_BS, 0, s>>>(flag, n, per, S.bcnt, keypart, gain, S.keys,
                                             S.vals);
  // every vertex may be listed: sort n pairs (the unlisted tail holds stale
  // keys, so the count is read first)
  i64 cnt = 0;
  (void)hipMemcpyAsync(&cnt, S.total, sizeof(i64), hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  k_bal_clear_u8<<<bl_grid(n, BAL_BS, 1u << 16), BAL_BS, 0, s>>>(keep, n);
  if (cnt <= 0) return;
  int end_bit = 32;
  while ((1ll << (end_bit - 32)) < (i64)k) end_bit++;
  size_t tb = (size_t)S.sort_bytes;
  hipcub::DeviceRadixSort::SortPairs(S.sort_tmp, tb, S.keys, S.keys2, S.vals, S.vals2, (int)cnt,
                                     0, end_bit, s);
  k_bal_fill_i64<<<1, 256, 0, s>>>(S.start, k, cnt);
  k_bal_starts<<<bl_grid(cnt, BAL_BS, 1u << 16), BAL_BS, 0, s>>>(S.keys2, cnt, S.start);
  k_bal_keep<<<bl_grid(cnt, BAL_BS, 1u << 16), BAL_BS, 0, s>>>(S.keys2, S.vals2, cnt, S.start,
                                                               limit, keep);
}

void launch_bal_avail(const i32* pc, i32 k, i64 cap, i64* avail, hipStream_t s) {
  k_bal_avail<<<1, 256, 0, s>>>(pc, k, cap, avail);
}

void launch_bal_apply(const u8* keep, i64 n, const i32* target, i32* part, i32* pc, u64* moved,
                      hipStream_t s) {
  k_bal_apply<<<bl_grid(n, BAL_BS, 1u << 16), BAL_BS, 0, s>>>(keep, n, target, part, pc, moved);
}

}  // namespace tgp
