// tgp_kernels.hip -- gfx950 kernels of the multilevel k-way partitioner.
//
// Every kernel reproduces one step of src/gpu_part.cu (iamrupamroyy/
// TheGraphPartitioning) under the SEQ schedule (each reference launch behaves
// as if its threads ran one at a time in ascending index), so results are
// bit-identical to the CPU oracle.  Layout and mapping are CDNA-first: 64-lane
// waves, one wave per vertex wherever a row is scanned, LDS tallies and
// histograms, integer atomics only where the result is order-independent.
//
// Floating point: compiled with -ffp-contract=off; every fp32 expression keeps
// the reference's operation order.  Per-part gain tallies are accumulated as
// integers when the row's sum of |w| is at most 2^24 (then every fp32 partial
// sum in row order is exact and equals the integer), otherwise one lane adds
// them in row order in fp32.
#include "tgp_internal.h"

#include <climits>

namespace tgp {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
__device__ __forceinline__ u64 lanes_below(int lane) { return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane)); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ u64 wave_max_u64(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const u64 t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}
__device__ __forceinline__ u64 wave_or_u64(u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

inline unsigned grid_for(i64 n, int per_block, unsigned cap = 1u << 20) {
  i64 b = (n + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > (i64)cap) b = cap;
  return (unsigned)b;
}

__device__ __forceinline__ i64 gtid() { return (i64)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ i64 gstride() { return (i64)gridDim.x * blockDim.x; }
// global wave index / count for kernels whose blocks hold whole waves
__device__ __forceinline__ i64 gwave() { return gtid() >> 6; }
__device__ __forceinline__ i64 nwaves() { return gstride() >> 6; }

__device__ __forceinline__ u32 hash32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ u64 hash64(u64 x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// ------------------------------------------------------------------ generic --
template <typename T>
__global__ void k_fill(T* p, i64 n, T v) {
  for (i64 i = gtid(); i < n; i += gstride()) p[i] = v;
}

constexpr int SCAN_BS = 256;
constexpr int SCAN_IPT = 8;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IPT;

// Exclusive scan of one tile of 2048 int64 values; bsum[b] = tile total.
__global__ __launch_bounds__(SCAN_BS) void k_scan_tile(const i64* in, i64* out, i64 n,
                                                       i64* bsum) {
  __shared__ i64 sw[SCAN_BS / WAVE];
  const i64 base = (i64)blockIdx.x * SCAN_TILE + (i64)threadIdx.x * SCAN_IPT;
  i64 v[SCAN_IPT];
  i64 s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_IPT; j++) {
    const i64 idx = base + j;
    v[j] = idx < n ? in[idx] : 0;
    s += v[j];
  }
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  i64 x = s;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const i64 t = __shfl_up(x, o);
    if (lane >= o) x += t;
  }
  if (lane == WAVE - 1) sw[wv] = x;
  __syncthreads();
  i64 woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < SCAN_BS / WAVE; w++) {
    if (w < wv) woff += sw[w];
    tot += sw[w];
  }
  i64 run = woff + x - s;
#pragma unroll
  for (int j = 0; j < SCAN_IPT; j++) {
    const i64 idx = base + j;
    if (idx < n) out[idx] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_add(i64* out, i64 n, const i64* boff) {
  const i64 add = boff[blockIdx.x];
  const i64 base = (i64)blockIdx.x * SCAN_TILE;
  for (int j = threadIdx.x; j < SCAN_TILE; j += SCAN_BS) {
    const i64 idx = base + j;
    if (idx < n) out[idx] += add;
  }
}

__global__ void k_copy1(const i64* src, i64* dst) { *dst = *src; }

__global__ __launch_bounds__(256) void k_level_census(DGraph g, i64 nedges, i64 heavy_deg,
                                                      i32* heavy_list, i32* heavy_count,
                                                      i32* nonunit) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > heavy_deg) heavy_list[atomicAdd(heavy_count, 1)] = (i32)v;
  }
  bool nu = false;
  for (i64 i = gtid(); i < g.m; i += gstride()) nu |= (g.w[i] != 1);
  if (nu) atomicOr(nonunit, 1);
}

// ----------------------------------------------------------------- matching --
// Exact parallel form of the sequential greedy matching of gpu_part.cu:55-78
// (thread t, in ascending order, takes its heaviest free neighbour, first
// position winning ties).  Rounds of deterministic reservations: every pending
// vertex s reserves all its free neighbours x with key (round, -s) and
// proposes its best free neighbour; t commits t-u when no pending s < t
// neighbours t or u and u > t or u is not pending.  The decisions equal the
// sequential ones because a commit happens only once every earlier vertex that
// could touch t or u has decided.
constexpr int MATCH_BS = 256;

// Only undecided vertices below t can block t, so each round evaluates the
// window [lo, lo + W) of vertex ids, lo = smallest undecided id: vertices
// above the window decide in later rounds with the same outcome (conditions
// only delay a decision, never change it) and the smallest undecided vertex
// always decides.  Every block recomputes lo from the stored value (the
// decided flags do not change inside the kernel) and block 0 stores it.
__device__ __forceinline__ i64 block_first_undecided(const u8* dec, i64 lo, i64 n) {
  __shared__ i64 s_found;
  for (i64 base = lo; base < n; base += blockDim.x) {
    if (threadIdx.x == 0) s_found = LLONG_MAX;
    __syncthreads();
    const i64 i = base + threadIdx.x;
    if (i < n && !dec[i]) atomicMin((unsigned long long*)&s_found, (unsigned long long)i);
    __syncthreads();
    const i64 f = s_found;
    __syncthreads();
    if (f != LLONG_MAX) return f;
  }
  return n;
}

__global__ __launch_bounds__(MATCH_BS) void k_match_reserve(DGraph g, const i32* mate,
                                                            const u8* dec, i32* pend, u64* R,
                                                            i32* prop, i64* lo_io, i64 W,
                                                            u32 round) {
  const i64 lo = block_first_undecided(dec, *lo_io, g.n);
  if (blockIdx.x == 0 && threadIdx.x == 0) *lo_io = lo;
  const int lane = lane_id();
  const u64 mykey_hi = (u64)round << 32;
  const i64 hi = lo + W < g.n ? lo + W : g.n;
  for (i64 base = lo + gwave() * WAVE; base < hi; base += nwaves() * WAVE) {
    const i64 t0 = base + lane;
    const bool alive = t0 < hi && !dec[t0];
    u64 am = __ballot(alive);
    if (alive) pend[t0] = (i32)round;
    // every undecided window vertex reserves its free neighbours and proposes
    // its heaviest free neighbour; rows are scanned by the whole wave
    while (am) {
      const int l = __ffsll((unsigned long long)am) - 1;
      const i32 t = (i32)(base + l);
      const i64 s0 = g.xadj[t], s1 = g.xadj[t + 1];
      const u64 key = mykey_hi | (u64)(0xFFFFFFFFu - (u32)t);
      u64 best = 0;
      for (i64 i = s0 + lane; i < s1; i += WAVE) {
        const i32 u = g.adj[i];
        if (u < g.n && u != t && mate[u] == -1) {
          atomicMax((unsigned long long*)&R[u], (unsigned long long)key);
          const i32 w = g.w[i];
          if (w >= 0) {  // gpu_part.cu:62,65: weight > maxWeight with maxWeight = -1
            const u64 c = ((u64)(u32)w << 32) | (u64)(0xFFFFFFFFu - (u32)(i - s0));
            best = c > best ? c : best;
          }
        }
      }
      best = wave_max_u64(best);
      if (lane == l) prop[t] = best ? g.adj[s0 + (i64)(0xFFFFFFFFu - (u32)best)] : -1;
      am &= am - 1;
    }
  }
}

// Decide and apply: reads only R, prop and pend, which this kernel does not
// write, so the decisions see the state at the start of the round.
__global__ __launch_bounds__(MATCH_BS) void k_match_decide(const u64* R, const i32* prop,
                                                           const i32* pend, i32* mate, u8* dec,
                                                           const i64* lo_io, i64 W, i64 n,
                                                           u32 round) {
  const i64 lo = *lo_io;
  const i64 hi = lo + W < n ? lo + W : n;
  for (i64 t = lo + gtid(); t < hi; t += gstride()) {
    if (pend[t] != (i32)round) continue;
    const u64 rt = R[t];
    const bool A = !((u32)(rt >> 32) == round && (i64)(0xFFFFFFFFu - (u32)rt) < t);
    const i32 u = prop[t];
    if (u < 0) {
      if (A) dec[t] = 1;
      continue;
    }
    const u64 ru = R[u];
    const bool C = !((u32)(ru >> 32) == round && (i64)(0xFFFFFFFFu - (u32)ru) < t);
    // u outside the window lies above t; u below lo is decided
    const bool B = (u > t) || (pend[u] != (i32)round);
    if (A && B && C) {
      mate[t] = u;
      mate[u] = (i32)t;
      dec[t] = 1;
      dec[u] = 1;
    }
  }
}

__global__ void k_match_init(i64 n, i32* mate, u8* dec, i32* pend, u64* R) {
  for (i64 i = gtid(); i < n; i += gstride()) {
    mate[i] = -1;
    dec[i] = 0;
    pend[i] = 0;
    R[i] = 0;
  }
}

// Symmetric levels (every arc s->x has its reverse x->s): the vertices that
// reserve x in a round are exactly x's undecided neighbours, so the round rule
// above reads: t may decide once no undecided neighbour lies below t
// (condition A), and may take its proposal u once no undecided neighbour of u
// lies below t (condition C).  Instead of recomputing these every round, a
// waiting vertex remembers one blocker, the LARGEST undecided vertex that
// violates the condition, and is re-examined only after that blocker has
// decided (decisions are permanent and the undecided set only shrinks, so a
// condition found true stays true).  Lower neighbours tend to decide in
// ascending order, so one rescan usually clears the condition.  The proposal
// (heaviest free neighbour, first position) is cached while it stays free.
constexpr i32 BLK_UNKNOWN = -1;
constexpr i32 BLK_NONE = -2;

// Row scans issue SCAN_U independent loads per lane before using any of them,
// so a long row costs about len/(64*SCAN_U) memory round trips.
constexpr int SCAN_U = 8;

// largest undecided neighbour u of x with u < lim, u != x (BLK_NONE if none)
__device__ __forceinline__ i32 wave_max_pending_below(const DGraph& g, i32 x, i32 lim,
                                                      const u8* dec, int lane) {
  const i64 s0 = g.xadj[x], s1 = g.xadj[x + 1];
  i32 best = BLK_NONE;
  for (i64 base = s0; base < s1; base += WAVE * SCAN_U) {
    i32 u[SCAN_U];
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) {
      const i64 i = base + j * WAVE + lane;
      u[j] = i < s1 ? g.adj[i] : INT_MAX;
    }
    u8 d[SCAN_U];
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) d[j] = (u[j] < lim && u[j] != x) ? dec[u[j]] : (u8)1;
#pragma unroll
    for (int j = 0; j < SCAN_U; j++)
      if (!d[j]) best = u[j] > best ? u[j] : best;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const i32 t = __shfl_xor(best, o);
    best = t > best ? t : best;
  }
  return best;
}

__device__ __forceinline__ i32 wave_best_free(const DGraph& g, i32 t, const i32* mate, int lane) {
  const i64 s0 = g.xadj[t], s1 = g.xadj[t + 1];
  u64 best = 0;
  for (i64 base = s0; base < s1; base += WAVE * SCAN_U) {
    i32 u[SCAN_U], w[SCAN_U], mt[SCAN_U];
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) {
      const i64 i = base + j * WAVE + lane;
      u[j] = i < s1 ? g.adj[i] : -1;
      w[j] = i < s1 ? g.w[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < SCAN_U; j++)
      mt[j] = (u[j] >= 0 && u[j] < g.n && u[j] != t && w[j] >= 0) ? mate[u[j]] : 0;
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) {
      if (mt[j] == -1) {
        const i64 i = base + j * WAVE + lane;
        const u64 c = ((u64)(u32)w[j] << 32) | (u64)(0xFFFFFFFFu - (u32)(i - s0));
        best = c > best ? c : best;
      }
    }
  }
  best = wave_max_u64(best);
  return best ? g.adj[s0 + (i64)(0xFFFFFFFFu - (u32)best)] : -1;
}

// decision codes
constexpr i32 DEC_WAIT = -2;
constexpr i32 DEC_EXHAUST = -1;

__device__ __forceinline__ bool blocker_stale(i32 b, const u8* dec) {
  return b == BLK_UNKNOWN || (b >= 0 && dec[b]);
}

// One lane per window vertex; lanes whose blocker or proposal must be
// refreshed are served one at a time by a cooperative scan of the whole wave.
// Rows longer than MATCH_HUGE arcs are not scanned here: the lane waits this
// round and queues a request that k_match_huge serves with a whole block
// after the round (its results are valid in the later state too).  Every
// cache is private to its vertex, and all reads of mate/dec see the state at
// the start of the round.
constexpr i64 MATCH_HUGE = 4096;
constexpr int SYM_VPW = 16;  // window vertices per wave (the rest of the lanes help scans)
enum { REQ_A = 0, REQ_P = 1, REQ_C = 2 };

__device__ __forceinline__ void push_request(int need_huge_lane, int lane, i32 kind, i32 t,
                                             i32* req, i32* nreq) {
  const u64 hm = __ballot(need_huge_lane);
  if (!hm) return;
  i32 ob = 0;
  if (lane == 0) ob = atomicAdd(nreq, (i32)__popcll(hm));
  ob = __shfl(ob, 0);
  if (need_huge_lane) {
    const i32 o = ob + __popcll(hm & lanes_below(lane));
    req[2 * o] = kind;
    req[2 * o + 1] = t;
  }
}

__device__ __forceinline__ i64 row_len(const DGraph& g, i32 x) {
  return g.xadj[x + 1] - g.xadj[x];
}

__global__ __launch_bounds__(MATCH_BS) void k_match_sym_decide(DGraph g, const i32* mate,
                                                               const u8* dec, i32* blkA,
                                                               i32* blkC, i32* prop,
                                                               i32* decision, i64* lo_io,
                                                               i64 W, i32* req, i32* nreq,
                                                               i32* nreq_next, int par) {
  // lo_io[0]: window low end of this round; lo_io[1 + parity]: where the scan
  // for the smallest undecided id starts.  This round reads slot `par` and
  // initialises the other one, which the apply kernel lowers to the smallest
  // id it leaves undecided and the next round reads.
  const i64 lo = block_first_undecided(dec, lo_io[1 + par], g.n);
  const i64 hi = lo + W < g.n ? lo + W : g.n;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    lo_io[0] = lo;
    lo_io[2 - par] = hi;
    *nreq_next = 0;  // the counter the next round appends to (read last round, done)
  }
  const int lane = lane_id();
  for (i64 base = lo + gwave() * SYM_VPW; base < hi; base += nwaves() * SYM_VPW) {
    const i64 e = base + lane;
    const bool alive = lane < SYM_VPW && e < hi && !dec[e];
    const i32 t = alive ? (i32)e : -1;
    if (!__ballot(alive)) continue;

    // condition A
    i32 a = alive ? blkA[t] : BLK_NONE;
    bool stale = alive && blocker_stale(a, dec);
    bool big = stale && row_len(g, t) > MATCH_HUGE;
    push_request(big, lane, REQ_A, t, req, nreq);
    u64 need = __ballot(stale && !big);
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(t, l);
      const i32 r = wave_max_pending_below(g, x, x, dec, lane);
      if (lane == l) {
        a = r;
        blkA[x] = r;
      }
      need &= need - 1;
    }
    const bool okA = alive && !big && a == BLK_NONE;
    // proposal
    i32 u = okA ? prop[t] : -1;
    stale = okA && (u == BLK_UNKNOWN || (u >= 0 && mate[u] != -1));
    big = stale && row_len(g, t) > MATCH_HUGE;
    push_request(big, lane, REQ_P, t, req, nreq);
    need = __ballot(stale && !big);
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(t, l);
      const i32 r = wave_best_free(g, x, mate, lane);
      if (lane == l) {
        u = r;
        prop[x] = r;
        blkC[x] = BLK_UNKNOWN;
      }
      need &= need - 1;
    }
    const bool okP = okA && !big;
    // condition B (u above t or already decided), then C
    const bool tryC = okP && u >= 0 && (u > t || dec[u]);
    i32 b = tryC ? blkC[t] : BLK_NONE;
    stale = tryC && blocker_stale(b, dec);
    big = stale && row_len(g, u) > MATCH_HUGE;
    push_request(big, lane, REQ_C, t, req, nreq);
    need = __ballot(stale && !big);
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(u, l);
      const i32 lim = __shfl(t, l);
      const i32 r = wave_max_pending_below(g, x, lim, dec, lane);
      if (lane == l) {
        b = r;
        blkC[lim] = r;
      }
      need &= need - 1;
    }
    if (alive) {
      i32 d = DEC_WAIT;
      if (okP) {
        if (u < 0) d = DEC_EXHAUST;
        else if (tryC && !big && b == BLK_NONE) d = u;
      }
      decision[t] = d;
    }
  }
}

// Serves queued scans of long rows with one block per request; runs after
// the round's apply, so it sees a later (still consistent) state.
__global__ __launch_bounds__(1024) void k_match_huge(DGraph g, const i32* mate, const u8* dec,
                                                     i32* blkA, i32* blkC, i32* prop,
                                                     const i32* req, const i32* nreq) {
  __shared__ u64 s_red[1024 / WAVE];
  const i32 cnt = *nreq;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i32 r = blockIdx.x; r < cnt; r += gridDim.x) {
    const i32 kind = req[2 * r], t = req[2 * r + 1];
    if (dec[t]) continue;
    i32 x = t, lim = t;
    if (kind == REQ_C) {
      x = prop[t];
      if (x < 0 || mate[x] != -1) continue;  // proposal gone: re-proposed next round
    }
    const i64 s0 = g.xadj[x], s1 = g.xadj[x + 1];
    u64 best = 0;  // REQ_A/C: 1 + max pending below lim; REQ_P: (w, ~pos) key
    for (i64 base = s0; base < s1; base += 1024 * 4) {
      i32 u[4], w[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const i64 i = base + j * 1024 + threadIdx.x;
        u[j] = i < s1 ? g.adj[i] : -1;
        w[j] = (kind == REQ_P && i < s1) ? g.w[i] : -1;
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const i64 i = base + j * 1024 + threadIdx.x;
        if (kind == REQ_P) {
          if (u[j] >= 0 && u[j] < g.n && u[j] != x && w[j] >= 0 && mate[u[j]] == -1) {
            const u64 c = ((u64)(u32)w[j] << 32) | (u64)(0xFFFFFFFFu - (u32)(i - s0));
            best = c > best ? c : best;
          }
        } else if (u[j] >= 0 && u[j] < lim && u[j] != x && !dec[u[j]]) {
          const u64 c = (u64)u[j] + 1;
          best = c > best ? c : best;
        }
      }
    }
    best = wave_max_u64(best);
    if (lane == 0) s_red[wv] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w2 = 1; w2 < 1024 / WAVE; w2++) best = s_red[w2] > best ? s_red[w2] : best;
      if (kind == REQ_A) {
        blkA[t] = best ? (i32)(best - 1) : BLK_NONE;
      } else if (kind == REQ_C) {
        blkC[t] = best ? (i32)(best - 1) : BLK_NONE;
      } else {
        prop[t] = best ? g.adj[s0 + (i64)(0xFFFFFFFFu - (u32)best)] : -1;
        blkC[t] = BLK_UNKNOWN;
      }
    }
    __syncthreads();
  }
}

// Applies the decisions of the window's vertices that were undecided when the
// round began (a vertex matched as a partner in this round had decided WAIT).
__global__ __launch_bounds__(MATCH_BS) void k_match_sym_apply(const i32* decision, i32* mate,
                                                              u8* dec, i64* lo_io, i64 W, i64 n,
                                                              int par) {
  const i64 lo = lo_io[0];
  const i64 hi = lo + W < n ? lo + W : n;
  i64 first = LLONG_MAX;
  for (i64 t = lo + gtid(); t < hi; t += gstride()) {
    if (dec[t]) continue;
    const i32 d = decision[t];
    if (d == DEC_EXHAUST) {
      dec[t] = 1;
    } else if (d >= 0) {
      mate[t] = d;
      mate[d] = (i32)t;
      dec[t] = 1;
      dec[d] = 1;
    } else if (t < first) {
      first = t;  // may since have been matched as a partner: the next scan skips it
    }
  }
  if (first != LLONG_MAX)
    atomicMin((unsigned long long*)&lo_io[2 - par], (unsigned long long)first);
}

__global__ void k_match_sym_init(i64 n, i32* mate, u8* dec, i32* blkA, i32* blkC, i32* prop) {
  for (i64 i = gtid(); i < n; i += gstride()) {
    mate[i] = -1;
    dec[i] = 0;
    blkA[i] = BLK_UNKNOWN;
    blkC[i] = BLK_UNKNOWN;
    prop[i] = BLK_UNKNOWN;
  }
}

// Symmetry test of a level: every arc s->x (s != x, x < n) must have x->s.
__global__ __launch_bounds__(256) void k_sym_insert(DGraph g, u64* table, i64 cap) {
  const u64 mask = (u64)cap - 1;
  const int lane = lane_id();
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < 0 || u >= g.n || u == v) continue;
      const u64 key = ((u64)v << 32) | (u64)(u32)u;
      u64 h = hash64(key) & mask;
      while (true) {
        const u64 old = atomicCAS((unsigned long long*)&table[h], ~0ull, (unsigned long long)key);
        if (old == ~0ull || old == key) break;
        h = (h + 1) & mask;
      }
    }
  }
}
__global__ __launch_bounds__(256) void k_sym_check(DGraph g, const u64* table, i64 cap,
                                                   i32* missing) {
  const u64 mask = (u64)cap - 1;
  const int lane = lane_id();
  bool miss = false;
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < 0 || u >= g.n || u == v) continue;
      const u64 key = ((u64)(u32)u << 32) | (u64)v;
      u64 h = hash64(key) & mask;
      while (true) {
        const u64 cur = table[h];
        if (cur == key) break;
        if (cur == ~0ull) {
          miss = true;
          break;
        }
        h = (h + 1) & mask;
      }
    }
  }
  if (__ballot(miss) && lane == 0) atomicOr(missing, 1);
}

// ------------------------------------------------------------------ relabel --
// gpu_part.cu:109-142 under SEQ: roots (unmatched, or the smaller end of a
// pair) get consecutive coarse ids in ascending order.
__global__ void k_root_flags(const i32* mate, i64 n, i64* f) {
  for (i64 t = gtid(); t < n; t += gstride()) {
    const i32 m = mate[t];
    f[t] = (m == -1 || t < m) ? 1 : 0;
  }
}
__global__ void k_relabel(const i32* mate, const i64* pos, i64 n, i32* new_ids, i32* tb) {
  for (i64 t = gtid(); t < n; t += gstride()) {
    const i32 m = mate[t];
    if (!(m == -1 || t < m)) continue;
    const i64 id = pos[t];
    new_ids[t] = (i32)id;
    tb[2 * id] = (i32)t;
    tb[2 * id + 1] = m;
    if (m != -1) new_ids[m] = (i32)id;
  }
}

// ------------------------------------------------------ reference contraction --
// gpu_part.cu:145-242: coarse row r (r < n_c) is built from FINE row r, then
// fine row match[r] when match[r] < n_c; x = new_ids[col], x == new_ids[r]
// skipped; bucket x%64 belongs to its first entry (claim order = output
// order), later entries equal to the owner add their weight, others drop.
// One wave per row walks the entries in order, 64 at a time; inside a chunk
// the first lane of each unclaimed bucket claims it (LDS atomicMin).
__global__ __launch_bounds__(WAVE) void k_contract_ref(DGraph g, const i32* mate,
                                                       const i32* new_ids, i64 n_c, i32* sadj,
                                                       i32* sw, i64* cnt) {
  __shared__ i32 s_first[WAVE], s_owner[WAVE], s_slot[WAVE], s_w[WAVE];
  const int lane = threadIdx.x;
  for (i64 r = blockIdx.x; r < n_c; r += gridDim.x) {
    const i32 me = new_ids[r];
    const i64 a0 = g.xadj[r], a1 = g.xadj[r + 1];
    const i32 mr = mate[r];
    i64 b0 = 0, b1 = 0;
    if (mr != -1 && mr < n_c) {
      b0 = g.xadj[mr];
      b1 = g.xadj[mr + 1];
    }
    const i64 la = a1 - a0, tot = la + (b1 - b0);
    u64 claimed = 0;
    int nclaimed = 0;
    s_owner[lane] = -1;
    s_w[lane] = 0;
    for (i64 base = 0; base < tot; base += WAVE) {
      const i64 p = base + lane;
      i32 x = -1, w = 0;
      bool valid = false;
      if (p < tot) {
        const i64 i = p < la ? a0 + p : b0 + (p - la);
        x = new_ids[g.adj[i]];
        w = g.w[i];
        valid = (x != me && x != -1);
      }
      const int b = valid ? (x & (WAVE - 1)) : 0;
      s_first[lane] = WAVE;
      __syncthreads();
      const bool fresh = valid && !((claimed >> b) & 1ull);
      if (fresh) atomicMin(&s_first[b], lane);
      __syncthreads();
      const bool claimer = fresh && s_first[b] == lane;
      const u64 cm = __ballot(claimer);
      if (claimer) {
        s_owner[b] = x;
        s_slot[b] = nclaimed + __popcll(cm & lanes_below(lane));
      }
      claimed |= wave_or_u64(claimer ? (1ull << b) : 0ull);
      nclaimed += __popcll(cm);
      __syncthreads();
      if (valid && s_owner[b] == x) atomicAdd(&s_w[b], w);
      __syncthreads();
    }
    if ((claimed >> lane) & 1ull) {
      const i64 o = r * WAVE + s_slot[lane];
      sadj[o] = s_owner[lane];
      sw[o] = s_w[lane];
    }
    if (lane == 0) cnt[r] = nclaimed;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_compact_rows(const i32* sadj, const i32* sw,
                                                      const i64* ox, i64 n_c, i32* oa,
                                                      i32* ow) {
  const int lane = lane_id();
  for (i64 r = gwave(); r < n_c; r += nwaves()) {
    const i64 o = ox[r], c = ox[r + 1] - o;
    if (lane < c) {
      oa[o + lane] = sadj[r * WAVE + lane];
      ow[o + lane] = sw[r * WAVE + lane];
    }
  }
}

// -------------------------------------------------------- repaired contraction --
// Coarse row c = entries of fine rows tb[2c] then tb[2c+1] (row order),
// relabelled, self loops dropped, each x kept at its first occurrence with the
// summed weight.  Per-row open-addressing tables (capacity pow2 >= 2d) record
// for each x its first position (atomicMin) and weight sum (atomicAdd).
__device__ __forceinline__ i64 next_pow2_i64(i64 x) {
  i64 p = 1;
  while (p < x) p <<= 1;
  return p;
}

__global__ void k_rep_sizes(const i64* xadj, const i32* tb, i64 n_c, i64* dsz, i64* csz) {
  for (i64 c = gtid(); c < n_c; c += gstride()) {
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    i64 d = xadj[r + 1] - xadj[r];
    if (mt >= 0) d += xadj[mt + 1] - xadj[mt];
    dsz[c] = d;
    csz[c] = d == 0 ? 0 : next_pow2_i64(2 * d);
  }
}

__device__ __forceinline__ void rep_entry(const DGraph& g, const i32* tb, i64 c, i64 p, i64 la,
                                          i64 a0, i64 b0, i32* col, i32* wt) {
  const i64 i = p < la ? a0 + p : b0 + (p - la);
  *col = g.adj[i];
  *wt = g.w[i];
}

__global__ __launch_bounds__(256) void k_rep_insert(DGraph g, const i32* tb, const i32* new_ids,
                                                    i64 n_c, const i64* toff, const i64* csz,
                                                    i32* tkey, i32* tpos, i32* tw) {
  const int lane = lane_id();
  for (i64 c = gwave(); c < n_c; c += nwaves()) {
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0, lb = mt >= 0 ? g.xadj[mt + 1] - b0 : 0;
    const i64 tot = la + lb;
    const u32 mask = (u32)(csz[c] - 1);
    i32* K = tkey + toff[c];
    i32* P = tpos + toff[c];
    i32* W = tw + toff[c];
    for (i64 p = lane; p < tot; p += WAVE) {
      i32 col, w;
      rep_entry(g, tb, c, p, la, a0, b0, &col, &w);
      const i32 x = new_ids[col];
      if (x == (i32)c || x == -1) continue;
      u32 h = hash32((u32)x) & mask;
      while (true) {
        const i32 old = atomicCAS(&K[h], -1, x);
        if (old == -1 || old == x) break;
        h = (h + 1) & mask;
      }
      atomicMin(&P[h], (i32)p);
      atomicAdd(&W[h], w);
    }
  }
}

__global__ __launch_bounds__(256) void k_rep_mark(DGraph g, const i32* tb, const i32* new_ids,
                                                  i64 n_c, const i64* toff, const i64* csz,
                                                  const i64* eoff, const i32* tkey,
                                                  const i32* tpos, i32* eslot, i64* cnt) {
  const int lane = lane_id();
  for (i64 c = gwave(); c < n_c; c += nwaves()) {
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0, lb = mt >= 0 ? g.xadj[mt + 1] - b0 : 0;
    const i64 tot = la + lb;
    const u32 mask = (u32)(csz[c] - 1);
    const i32* K = tkey + toff[c];
    const i32* P = tpos + toff[c];
    i32* E = eslot + eoff[c];
    i64 reps = 0;
    for (i64 p = lane; p < tot; p += WAVE) {
      i32 col, w;
      rep_entry(g, tb, c, p, la, a0, b0, &col, &w);
      const i32 x = new_ids[col];
      i32 s = -1;
      if (!(x == (i32)c || x == -1)) {
        u32 h = hash32((u32)x) & mask;
        while (K[h] != x) h = (h + 1) & mask;
        if (P[h] == (i32)p) s = (i32)h;
      }
      E[p] = s;
      reps += (s >= 0);
    }
    reps = wave_sum(reps);
    if (lane == 0) cnt[c] = reps;
  }
}

__global__ __launch_bounds__(256) void k_rep_emit(const i64* eoff, const i64* dsz,
                                                  const i32* eslot, const i64* toff,
                                                  const i32* tkey, const i32* tw, const i64* ox,
                                                  i64 n_c, i32* oa, i32* ow) {
  const int lane = lane_id();
  for (i64 c = gwave(); c < n_c; c += nwaves()) {
    const i64 tot = dsz[c];
    const i32* E = eslot + eoff[c];
    const i32* K = tkey + toff[c];
    const i32* W = tw + toff[c];
    i64 run = ox[c];
    for (i64 base = 0; base < tot; base += WAVE) {
      const i64 p = base + lane;
      const i32 s = p < tot ? E[p] : -1;
      const u64 bm = __ballot(s >= 0);
      if (s >= 0) {
        const i64 o = run + __popcll(bm & lanes_below(lane));
        oa[o] = K[s];
        ow[o] = W[s];
      }
      run += __popcll(bm);
    }
  }
}

// Unique undirected pairs of the coarsest graph (gpu_part.cu:810-827).
__global__ __launch_bounds__(256) void k_pairs(DGraph g, u64* table, i64 cap, u64* count) {
  const u64 mask = (u64)cap - 1;
  u64 local = 0;
  const int lane = lane_id();
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : g.m;
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u >= 0 && u < g.n && u != v) {
        const u64 a = (u64)(v < u ? v : u), b = (u64)(v < u ? u : v);
        const u64 key = (a << 32) | b;
        u64 h = hash64(key) & mask;
        while (true) {
          const u64 old = atomicCAS((unsigned long long*)&table[h], ~0ull, (unsigned long long)key);
          if (old == ~0ull) {
            local++;
            break;
          }
          if (old == key) break;
          h = (h + 1) & mask;
        }
      }
    }
  }
  local = wave_sum(local);
  if (lane == 0 && local) atomicAdd((unsigned long long*)count, (unsigned long long)local);
}

// --------------------------------------------------------------- projection --
// gpu_part.cu:763-772 then the host recount 1244-1252 (-1 -> part 0).
__global__ void k_project(const i32* tb, const i32* part_c, i64 n_c, i32* part_f, i64 n_f,
                          i32 k) {
  for (i64 c = gtid(); c < n_c; c += gstride()) {
    const i32 cp = part_c[c];
    if (cp >= k || cp < 0) continue;
    const i32 v1 = tb[2 * c], v2 = tb[2 * c + 1];
    if (v1 >= 0 && v1 < n_f) part_f[v1] = cp;
    if (v2 >= 0 && v2 < n_f) part_f[v2] = cp;
  }
}

constexpr int MAXK_LDS = 4096;

__global__ __launch_bounds__(256) void k_sanitize_hist(i32* part, i64 n, i32 k, i32* pc) {
  __shared__ i32 h[MAXK_LDS];
  for (int q = threadIdx.x; q < k; q += blockDim.x) h[q] = 0;
  __syncthreads();
  for (i64 i = gtid(); i < n; i += gstride()) {
    i32 p = part[i];
    if (p < 0 || p >= k) {
      p = 0;
      part[i] = 0;
    }
    atomicAdd(&h[p], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < k; q += blockDim.x)
    if (h[q]) atomicAdd(&pc[q], h[q]);
}

__global__ void k_block_init(i32* part, i64 n, i32 k) {
  for (i64 i = gtid(); i < n; i += gstride()) part[i] = (i32)((i * (i64)k) / n);
}

// --------------------------------------------------------------- refinement --
// gpu_part.cu:286-295: the balance term depends only on (q, phase, whether
// the source part is oversized), so it is tabulated per pass with the
// reference's operation order: b = (c*(T-d))/T; [+A if src big]; [+B if dst small].
__global__ void k_bg_table(const i32* pc, i32 k, i64 n, int phase, float* bg, u8* src_big) {
  const i64 T = n / k;
  const float Tf = (float)T;
  for (i64 q = gtid(); q < k; q += gstride()) {
    const float d = (float)pc[q];
    const float base = phase ? 500.0f * (Tf - d) / Tf : 200.0f * (Tf - d) / Tf;
    const float A = phase ? 250.0f : 100.0f;
    const float B = phase ? 125.0f : 50.0f;
    const bool small = d < Tf * 0.95f;
    float v0 = base;
    if (small) v0 = v0 + B;
    float v1 = base + A;
    if (small) v1 = v1 + B;
    bg[q] = v0;
    bg[k + q] = v1;
    src_big[q] = d > Tf * 1.05f;
  }
}

// gpu_part.cu:258-312: tallies, best target, candidate flag.  One wave per
// vertex; LDS tally of k ints (or floats in the exact-order fallback).
template <bool UNIT, bool LDS>
__global__ __launch_bounds__(WAVE) void k_gain(DGraph g, i64 nedges, const i32* part, i32 k,
                                               const float* bg_g, const u8* src_big,
                                               float* gain, i32* target, u8* cand, int phase,
                                               i32* gscratch, const i32* list_a, i64 cnt_a,
                                               const i32* list_b, const i32* cnt_b) {
  extern __shared__ i32 lds[];
  const int lane = threadIdx.x;
  i32* tally;
  const float* bg;
  if (LDS) {
    float* sbg = (float*)(lds + k);
    for (int q = lane; q < 2 * k; q += WAVE) sbg[q] = bg_g[q];
    tally = lds;
    bg = sbg;
  } else {
    tally = gscratch + (i64)blockIdx.x * k;
    bg = bg_g;
  }
  __syncthreads();
  float* ftally = (float*)tally;
  const i64 total = cnt_a + (list_b ? (i64)*cnt_b : 0);
  for (i64 e = blockIdx.x; e < total; e += gridDim.x) {
    const i64 v = e < cnt_a ? (list_a ? list_a[e] : e) : list_b[e - cnt_a];
    const i32 cur = part[v];
    if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
    for (int q = lane; q < k; q += WAVE) tally[q] = 0;
    __syncthreads();
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    i64 absw = 0;
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < g.n) {
        const i32 p = part[u];
        if (p < k && p >= 0) {
          const i32 w = UNIT ? 1 : g.w[i];
          atomicAdd(&tally[p], w);
          absw += w < 0 ? -(i64)w : (i64)w;
        }
      }
    }
    absw = UNIT ? 0 : wave_sum(absw);
    __syncthreads();
    // Integer tallies equal the fp32 row-order sums while no partial sum can
    // exceed 2^24: bounded by the row's sum of |w| (or its length for unit w).
    const bool exact_int = UNIT ? (s1 - s0) <= (i64)(1 << 24) : absw <= (i64)(1 << 24);
    if (!exact_int) {
      for (int q = lane; q < k; q += WAVE) ftally[q] = 0.0f;
      __syncthreads();
      if (lane == 0) {
        for (i64 i = s0; i < s1; i++) {
          const i32 u = g.adj[i];
          if (u < g.n) {
            const i32 p = part[u];
            if (p < k && p >= 0) ftally[p] += (float)g.w[i];
          }
        }
      }
      __syncthreads();
    }
    const float ccur = exact_int ? (float)tally[cur] : ftally[cur];
    const float* bgr = bg + (src_big[cur] ? k : 0);
    float bgv = -INFINITY;
    i32 bq = INT_MAX;
    for (int q = lane; q < k; q += WAVE) {
      if (q == cur) continue;
      const float cq = exact_int ? (float)tally[q] : ftally[q];
      const float eg = cq - ccur;
      const float gg = eg + bgr[q];
      if (gg == gg && (gg > bgv || (gg == bgv && q < bq))) {
        bgv = gg;
        bq = q;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float og = __shfl_xor(bgv, o);
      const i32 oq = __shfl_xor(bq, o);
      if (og > bgv || (og == bgv && oq < bq)) {
        bgv = og;
        bq = oq;
      }
    }
    if (lane == 0) {
      float best;
      i32 tgt;
      if (bq == INT_MAX || bgv < -1e9f) {
        best = -1e9f;
        tgt = cur;
      } else if (bgv == -1e9f) {
        best = -1e9f;
        tgt = bq < cur ? bq : cur;
      } else {
        best = bgv;
        tgt = bq;
      }
      gain[v] = best;
      target[v] = tgt;
      cand[v] = phase ? (best > -50.0f) : (best > -10.0f);  // gpu_part.cu:308-312
    }
    __syncthreads();
  }
}

// Per-pass order of the parts by (balance term desc, part asc) for both
// source-size variants, and for every position the end of its run of equal
// balance terms.  For a vertex whose tally of part q is zero the gain is
// (0 - c_cur) + b_q, which is non-decreasing in b_q, so its best such part is
// found by walking runs from the top (first allowed part of a run = lowest q).
__global__ __launch_bounds__(1024) void k_bg_order(const float* bg, i32 k, i32* ord,
                                                   i32* run_end) {
  __shared__ u64 key[MAXK_LDS];
  __shared__ i32 rend[MAXK_LDS];
  i32 P = 1;
  while (P < k) P <<= 1;
  for (int s = 0; s < 2; s++) {
    const float* b = bg + s * k;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
      if (i < k) {
        const u32 bits = __float_as_uint(b[i] + 0.0f);  // -0 -> +0
        const u32 ordk = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);  // ascending
        key[i] = ((u64)(~ordk) << 32) | (u32)i;  // descending balance term, then ascending q
      } else {
        key[i] = ~0ull;
      }
    }
    __syncthreads();
    for (i32 kk = 2; kk <= P; kk <<= 1) {
      for (i32 j = kk >> 1; j > 0; j >>= 1) {
        for (i32 t = threadIdx.x; t < P / 2; t += blockDim.x) {
          const i32 lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
          const i32 hi = lo + j;
          const bool asc = (lo & kk) == 0;
          const u64 a = key[lo], c = key[hi];
          if ((a > c) == asc) {
            key[lo] = c;
            key[hi] = a;
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < k; i += blockDim.x) {
      const i32 q = (i32)(key[i] & 0xffffffffu);
      ord[s * k + i] = q;
      const bool last = (i == k - 1) || (b[(i32)(key[i + 1] & 0xffffffffu)] != b[q]);
      rend[i] = last ? i + 1 : INT_MAX;
    }
    __syncthreads();
    for (i32 d = 1; d < k; d <<= 1) {
      i32 v[4];
      int c = 0;
      for (int i = threadIdx.x; i < k; i += blockDim.x, c++) v[c] = (i + d < k) ? rend[i + d] : INT_MAX;
      __syncthreads();
      c = 0;
      for (int i = threadIdx.x; i < k; i += blockDim.x, c++) rend[i] = v[c] < rend[i] ? v[c] : rend[i];
      __syncthreads();
    }
    for (int i = threadIdx.x; i < k; i += blockDim.x) run_end[s * k + i] = rend[i];
    __syncthreads();
  }
}

// One thread per vertex whose range has at most GAIN_LIGHT arcs: the tallies
// of up to GAIN_SLOTS distinct parts live in registers and are accumulated
// in row order (exactly the reference's fp32 sums).  Vertices with more
// distinct parts are deferred to the wave kernel.
constexpr int GAIN_SLOTS = 16;
constexpr int GAIN_LIGHT = 64;

template <bool UNIT>
__global__ __launch_bounds__(256) void k_gain_light(DGraph g, i64 nedges, const i32* part, i32 k,
                                                    const float* bg, const u8* src_big,
                                                    const i32* ord, const i32* run_end,
                                                    float* gain, i32* target, u8* cand,
                                                    int phase, i32* defer, i32* ndefer) {
  const int lane = lane_id();
  for (i64 base = gtid() - lane; base < g.n; base += gstride()) {
    const i64 v = base + lane;
    bool mine = false, over = false;
    i32 cur = -1;
    i32 keys[GAIN_SLOTS];
    float vals[GAIN_SLOTS];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < GAIN_SLOTS; j++) {
      keys[j] = -1;
      vals[j] = 0.0f;
    }
    if (v < g.n) {
      const i64 s0 = g.xadj[v];
      const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
      cur = part[v];
      mine = (s1 - s0 <= GAIN_LIGHT) && cur >= 0 && cur < k;
      if (mine) {
        for (i64 i = s0; i < s1; i++) {
          const i32 u = g.adj[i];
          if (u >= g.n) continue;
          const i32 p = part[u];
          if (p < 0 || p >= k) continue;
          const float wt = UNIT ? 1.0f : (float)g.w[i];
          bool found = false;
#pragma unroll
          for (int j = 0; j < GAIN_SLOTS; j++) {
            if (keys[j] == p) {
              vals[j] += wt;
              found = true;
            }
          }
          if (!found) {
            if (cnt < GAIN_SLOTS) {
#pragma unroll
              for (int j = 0; j < GAIN_SLOTS; j++) {
                if (j == cnt) {
                  keys[j] = p;
                  vals[j] = wt;
                }
              }
              cnt++;
            } else {
              over = true;
              break;
            }
          }
        }
      }
    }
    // defer vertices with too many distinct parts (one atomic per wave)
    const u64 om = __ballot(mine && over);
    if (om) {
      i32 ob = 0;
      if (lane == 0) ob = atomicAdd(ndefer, (i32)__popcll(om));
      ob = __shfl(ob, 0);
      if (mine && over) defer[ob + __popcll(om & lanes_below(lane))] = (i32)v;
    }
    if (!mine || over) continue;
    float ccur = 0.0f;
#pragma unroll
    for (int j = 0; j < GAIN_SLOTS; j++)
      if (keys[j] == cur) ccur = vals[j];
    const int sv = src_big[cur] ? 1 : 0;
    const float* bgs = bg + sv * k;
    // best over the parts present in the row
    float bgv = -INFINITY;
    i32 bq = INT_MAX;
#pragma unroll
    for (int j = 0; j < GAIN_SLOTS; j++) {
      const i32 q = keys[j];
      if (q >= 0 && q != cur) {
        const float eg = vals[j] - ccur;
        const float gg = eg + bgs[q];
        if (gg > bgv || (gg == bgv && q < bq)) {
          bgv = gg;
          bq = q;
        }
      }
    }
    // best over the parts absent from the row (tally 0.0f)
    const float e0 = 0.0f - ccur;
    const i32* o = ord + sv * k;
    const i32* re = run_end + sv * k;
    bool found = false;
    float g1 = 0.0f;
    i32 q1 = INT_MAX;
    for (i32 i = 0; i < k;) {
      const i32 e = re[i];
      i32 j = i;
      i32 q = -1;
      for (; j < e; j++) {
        q = o[j];
        bool tch = (q == cur);
#pragma unroll
        for (int t = 0; t < GAIN_SLOTS; t++) tch |= (keys[t] == q);
        if (!tch) break;
      }
      if (j < e) {
        const float gg = e0 + bgs[q];
        if (!found) {
          found = true;
          g1 = gg;
          q1 = q;
        } else if (gg == g1) {
          q1 = q < q1 ? q : q1;
        } else {
          break;
        }
      }
      i = e;
    }
    if (found && (g1 > bgv || (g1 == bgv && q1 < bq))) {
      bgv = g1;
      bq = q1;
    }
    float best;
    i32 tgt;
    if (bq == INT_MAX || bgv < -1e9f) {
      best = -1e9f;
      tgt = cur;
    } else if (bgv == -1e9f) {
      best = -1e9f;
      tgt = bq < cur ? bq : cur;
    } else {
      best = bgv;
      tgt = bq;
    }
    gain[v] = best;
    target[v] = tgt;
    cand[v] = phase ? (best > -50.0f) : (best > -10.0f);
  }
}

// gpu_part.cu:258-312 with G lanes per vertex: each group keeps a dense LDS
// tally of k ints (exact while the row's sum of |w| <= 2^24, else one lane
// redoes the row in fp32 in row order), then every lane scores k/G parts and
// the group reduces by (gain desc, part asc); a zero tally scores exactly like
// an absent part, so the dense scan is the reference's loop over q.  G = 8
// for rows of at most 64 arcs (eight vertices per wave in flight), G = 64 for
// the heavy list.
template <int G, bool UNIT>
__global__ __launch_bounds__(256) void k_gain_group(DGraph g, i64 nedges, const i32* part, i32 k,
                                                    const float* bg_g, const u8* src_big,
                                                    float* gain, i32* target, u8* cand,
                                                    int phase, const i32* list, i64 cnt,
                                                    i64 light_max) {
  extern __shared__ i32 lds[];
  constexpr int GPB = 256 / G;  // groups per block
  float* sbg = (float*)lds;
  for (int q = threadIdx.x; q < 2 * k; q += blockDim.x) sbg[q] = bg_g[q];
  __syncthreads();
  const int gl = threadIdx.x & (G - 1);
  const int grp = threadIdx.x / G;
  i32* tally = lds + 2 * k + grp * k;
  float* ftally = (float*)tally;
  const i64 total = list ? cnt : g.n;
  const i64 ngroups = (i64)gridDim.x * GPB;
  for (i64 e = (i64)blockIdx.x * GPB + grp; e < total; e += ngroups) {
    const i64 v = list ? list[e] : e;
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const i32 cur = part[v];
    if (!list && s1 - s0 > light_max) continue;   // heavy rows: the G = 64 launch
    if (cur >= k || cur < 0) continue;            // gpu_part.cu:262
    for (int q = gl; q < k; q += G) tally[q] = 0;
    __builtin_amdgcn_wave_barrier();
    i64 absw = 0;
    for (i64 i = s0 + gl; i < s1; i += G) {
      const i32 u = g.adj[i];
      if (u < g.n) {
        const i32 p = part[u];
        if (p < k && p >= 0) {
          const i32 w = UNIT ? 1 : g.w[i];
          atomicAdd(&tally[p], w);
          if (!UNIT) absw += w < 0 ? -(i64)w : (i64)w;
        }
      }
    }
    if (!UNIT) {
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) absw += __shfl_xor(absw, o, G);
    }
    __builtin_amdgcn_wave_barrier();
    const bool exact_int = UNIT ? (s1 - s0) <= (i64)(1 << 24) : absw <= (i64)(1 << 24);
    if (!exact_int) {
      for (int q = gl; q < k; q += G) ftally[q] = 0.0f;
      __builtin_amdgcn_wave_barrier();
      if (gl == 0) {
        for (i64 i = s0; i < s1; i++) {
          const i32 u = g.adj[i];
          if (u < g.n) {
            const i32 p = part[u];
            if (p < k && p >= 0) ftally[p] += (float)g.w[i];
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    const float ccur = exact_int ? (float)tally[cur] : ftally[cur];
    const float* bgr = sbg + (src_big[cur] ? k : 0);
    float bgv = -INFINITY;
    i32 bq = INT_MAX;
    for (int q = gl; q < k; q += G) {
      if (q == cur) continue;
      const float cq = exact_int ? (float)tally[q] : ftally[q];
      const float eg = cq - ccur;
      const float gg = eg + bgr[q];
      if (gg > bgv || (gg == bgv && q < bq)) {
        bgv = gg;
        bq = q;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      const float og = __shfl_xor(bgv, o, G);
      const i32 oq = __shfl_xor(bq, o, G);
      if (og > bgv || (og == bgv && oq < bq)) {
        bgv = og;
        bq = oq;
      }
    }
    if (gl == 0) {
      float best;
      i32 tgt;
      if (bq == INT_MAX || bgv < -1e9f) {
        best = -1e9f;
        tgt = cur;
      } else if (bgv == -1e9f) {
        best = -1e9f;
        tgt = bq < cur ? bq : cur;
      } else {
        best = bgv;
        tgt = bq;
      }
      gain[v] = best;
      target[v] = tgt;
      cand[v] = phase ? (best > -50.0f) : (best > -10.0f);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Rows longer than GAIN_HUGE arcs: one block of 1024 threads per vertex with
// four independent loads per thread in flight; block-wide LDS tally and pick.
constexpr int GAIN_HUGE = 4096;
constexpr int HUGE_BS = 1024;

template <bool UNIT>
__global__ __launch_bounds__(HUGE_BS) void k_gain_huge(DGraph g, i64 nedges, const i32* part,
                                                       i32 k, const float* bg_g,
                                                       const u8* src_big, float* gain,
                                                       i32* target, u8* cand, int phase,
                                                       const i32* list, i64 cnt) {
  extern __shared__ i32 lds[];
  __shared__ i64 s_abs[HUGE_BS / WAVE];
  __shared__ float s_g[HUGE_BS / WAVE];
  __shared__ i32 s_q[HUGE_BS / WAVE];
  float* sbg = (float*)lds;
  i32* tally = lds + 2 * k;
  float* ftally = (float*)tally;
  for (int q = threadIdx.x; q < 2 * k; q += HUGE_BS) sbg[q] = bg_g[q];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 e = blockIdx.x; e < cnt; e += gridDim.x) {
    const i64 v = list[e];
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const i32 cur = part[v];
    if (cur >= k || cur < 0) continue;
    for (int q = threadIdx.x; q < k; q += HUGE_BS) tally[q] = 0;
    __syncthreads();
    i64 absw = 0;
    for (i64 base = s0; base < s1; base += 4 * HUGE_BS) {
      i32 u[4], w[4], p[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const i64 i = base + j * HUGE_BS + threadIdx.x;
        u[j] = i < s1 ? g.adj[i] : INT_MAX;
        w[j] = (UNIT || i >= s1) ? 1 : g.w[i];
      }
#pragma unroll
      for (int j = 0; j < 4; j++) p[j] = u[j] < g.n ? part[u[j]] : -1;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (p[j] >= 0 && p[j] < k) {
          atomicAdd(&tally[p[j]], w[j]);
          if (!UNIT) absw += w[j] < 0 ? -(i64)w[j] : (i64)w[j];
        }
      }
    }
    if (!UNIT) {
      absw = wave_sum(absw);
      if (lane == 0) s_abs[wv] = absw;
    }
    __syncthreads();
    i64 tot = 0;
    if (!UNIT)
      for (int w2 = 0; w2 < HUGE_BS / WAVE; w2++) tot += s_abs[w2];
    const bool exact_int = UNIT ? (s1 - s0) <= (i64)(1 << 24) : tot <= (i64)(1 << 24);
    if (!exact_int) {
      __syncthreads();
      for (int q = threadIdx.x; q < k; q += HUGE_BS) ftally[q] = 0.0f;
      __syncthreads();
      if (threadIdx.x == 0) {
        for (i64 i = s0; i < s1; i++) {
          const i32 uu = g.adj[i];
          if (uu < g.n) {
            const i32 pp = part[uu];
            if (pp < k && pp >= 0) ftally[pp] += (float)g.w[i];
          }
        }
      }
      __syncthreads();
    }
    const float ccur = exact_int ? (float)tally[cur] : ftally[cur];
    const float* bgr = sbg + (src_big[cur] ? k : 0);
    float bgv = -INFINITY;
    i32 bq = INT_MAX;
    for (int q = threadIdx.x; q < k; q += HUGE_BS) {
      if (q == cur) continue;
      const float cq = exact_int ? (float)tally[q] : ftally[q];
      const float eg = cq - ccur;
      const float gg = eg + bgr[q];
      if (gg > bgv || (gg == bgv && q < bq)) {
        bgv = gg;
        bq = q;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float og = __shfl_xor(bgv, o);
      const i32 oq = __shfl_xor(bq, o);
      if (og > bgv || (og == bgv && oq < bq)) {
        bgv = og;
        bq = oq;
      }
    }
    if (lane == 0) {
      s_g[wv] = bgv;
      s_q[wv] = bq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w2 = 1; w2 < HUGE_BS / WAVE; w2++) {
        if (s_g[w2] > bgv || (s_g[w2] == bgv && s_q[w2] < bq)) {
          bgv = s_g[w2];
          bq = s_q[w2];
        }
      }
      float best;
      i32 tgt;
      if (bq == INT_MAX || bgv < -1e9f) {
        best = -1e9f;
        tgt = cur;
      } else if (bgv == -1e9f) {
        best = -1e9f;
        tgt = bq < cur ? bq : cur;
      } else {
        best = bgv;
        tgt = bq;
      }
      gain[v] = best;
      target[v] = tgt;
      cand[v] = phase ? (best > -50.0f) : (best > -10.0f);
    }
    __syncthreads();
  }
}

// Heavy rows split into medium (64 < len <= 4096, one wave) and huge (> 4096,
// one block): entries of `heavy` are partitioned in place by the census.
__global__ void k_split_heavy(DGraph g, i64 nedges, const i32* heavy, i64 nheavy, i32* medium,
                              i32* nmedium, i32* huge, i32* nhuge) {
  for (i64 e = gtid(); e < nheavy; e += gstride()) {
    const i32 v = heavy[e];
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > GAIN_HUGE) huge[atomicAdd(nhuge, 1)] = v;
    else medium[atomicAdd(nmedium, 1)] = v;
  }
}

// gpu_part.cu:316-333 (+ the move-eligibility test of 337-343 folded in).
__device__ __forceinline__ bool move_ok(i32 src, i32 dst, i32 k) {
  return src < k && dst < k && src != dst && src >= 0 && dst >= 0;
}

__global__ __launch_bounds__(256) void k_token_thread(DGraph g, i64 nedges, const float* gain,
                                                      const u8* cand, const i32* part,
                                                      const i32* target, i32 k, i64* flag,
                                                      i64 heavy_deg) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > heavy_deg) continue;
    if (!cand[v]) {
      flag[v] = 0;
      continue;
    }
    const float gv = gain[v];
    const float gv10 = gv + 10.0f;
    bool win = true;
    for (i64 i = s0; i < s1; i++) {
      const i32 u = g.adj[i];
      if (u < g.n && cand[u]) {
        const float gu = gain[u];
        if (gu > gv10 || (gu == gv && u < v)) {
          win = false;
          break;
        }
      }
    }
    flag[v] = (win && move_ok(part[v], target[v], k)) ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void k_token_heavy(DGraph g, i64 nedges, const float* gain,
                                                     const u8* cand, const i32* part,
                                                     const i32* target, i32 k, i64* flag,
                                                     const i32* list, i64 cnt) {
  const int lane = lane_id();
  for (i64 e = gwave(); e < cnt; e += nwaves()) {
    const i32 v = list[e];
    if (!cand[v]) {
      if (lane == 0) flag[v] = 0;
      continue;
    }
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const float gv = gain[v];
    const float gv10 = gv + 10.0f;
    bool lose = false;
    for (i64 base = s0; base < s1; base += WAVE) {
      const i64 i = base + lane;
      bool l = false;
      if (i < s1) {
        const i32 u = g.adj[i];
        if (u < g.n && cand[u]) {
          const float gu = gain[u];
          l = (gu > gv10 || (gu == gv && u < v));
        }
      }
      if (__ballot(l)) {
        lose = true;
        break;
      }
    }
    if (lane == 0) flag[v] = (!lose && move_ok(part[v], target[v], k)) ? 1 : 0;
  }
}

// gpu_part.cu:337-350 under SEQ: winners in ascending vertex order.
__global__ __launch_bounds__(256) void k_scatter_moves(const i64* flag, const i64* pos,
                                                       const i32* part, const i32* target,
                                                       const float* gain, i64 n, i32 k,
                                                       i32* mv_v, i32* mv_src, i32* mv_dst,
                                                       float* mv_gain, i32* outc) {
  __shared__ i32 h[MAXK_LDS];
  for (int q = threadIdx.x; q < k; q += blockDim.x) h[q] = 0;
  __syncthreads();
  for (i64 v = gtid(); v < n; v += gstride()) {
    if (!flag[v]) continue;
    const i64 p = pos[v];
    const i32 src = part[v];
    mv_v[p] = (i32)v;
    mv_src[p] = src;
    mv_dst[p] = target[v];
    mv_gain[p] = gain[v];
    atomicAdd(&h[src], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < k; q += blockDim.x)
    if (h[q]) atomicAdd(&outc[q], h[q]);
}

// gpu_part.cu:357-369.
__global__ void k_capacity(const i32* pc, const i32* outc, i32 k, i64 n, int phase,
                           i64* avail) {
  const i64 T = n / k;
  const float Tf = (float)T;
  const i64 slack0 = n / ((i64)k * 10);
  const i64 slack = slack0 > 100 ? slack0 : 100;
  const i64 imb = (i64)(Tf * 0.05f);
  const i64 maxSize = T + imb + slack;
  const i64 minSize = T - imb - slack;
  for (i64 q = gtid(); q < k; q += gstride()) {
    const i64 current = pc[q];
    i64 a = T - current + outc[q] + slack;
    if (a < 0) a = 0;
    if (a > maxSize - current + slack) a = maxSize - current + slack;
    if (current <= minSize && !phase) a = a < slack / 2 ? a : slack / 2;
    avail[q] = a;
  }
}

// ------------------------------------------------------------ bitonic network --
// gpu_part.cu:472-506: the exact compare-exchange sequence on (gain, slot)
// pairs.  The element at the higher index of each pair compares itself with
// the lower one: swap = (gain[hi] < gain[lo]) ^ ((hi & kk) == 0); pairs with
// hi >= len are skipped.  Stages with stride < 2048 run inside LDS tiles.
constexpr int BT_ELEMS = 2048;
constexpr int BT_THREADS = 1024;

__global__ void k_sort_init(const float* mv_gain, i64 M, i64 len, float* key, i32* idx) {
  for (i64 i = gtid(); i < len; i += gstride()) {
    if (i < M) {
      key[i] = mv_gain[i];
      idx[i] = (i32)i;
    } else {
      key[i] = -1e9f;
      idx[i] = -1;
    }
  }
}

__device__ __forceinline__ void bt_stage_lds(float* sk, i32* si, i64 base, i64 len, i64 kk,
                                             int j) {
  const int t = threadIdx.x;
  const int lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
  const int hi = lo + j;
  if (base + hi < len) {
    const bool up = ((base + hi) & kk) == 0;
    const float a = sk[hi], b = sk[lo];
    if ((a < b) ^ up) {
      sk[hi] = b;
      sk[lo] = a;
      const i32 tt = si[hi];
      si[hi] = si[lo];
      si[lo] = tt;
    }
  }
}

__global__ __launch_bounds__(BT_THREADS) void k_bt_local(float* key, i32* idx, i64 len,
                                                         i64 kk_fixed) {
  __shared__ float sk[BT_ELEMS];
  __shared__ i32 si[BT_ELEMS];
  const i64 base = (i64)blockIdx.x * BT_ELEMS;
  for (int e = threadIdx.x; e < BT_ELEMS; e += BT_THREADS) {
    const i64 gi = base + e;
    if (gi < len) {
      sk[e] = key[gi];
      si[e] = idx[gi];
    }
  }
  __syncthreads();
  if (kk_fixed == 0) {
    for (i64 kk = 2; kk <= BT_ELEMS && kk <= len; kk <<= 1) {
      for (int j = (int)(kk >> 1); j > 0; j >>= 1) {
        bt_stage_lds(sk, si, base, len, kk, j);
        __syncthreads();
      }
    }
  } else {
    for (int j = BT_ELEMS >> 1; j > 0; j >>= 1) {
      bt_stage_lds(sk, si, base, len, kk_fixed, j);
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < BT_ELEMS; e += BT_THREADS) {
    const i64 gi = base + e;
    if (gi < len) {
      key[gi] = sk[e];
      idx[gi] = si[e];
    }
  }
}

__global__ __launch_bounds__(256) void k_bt_global(float* key, i32* idx, i64 len, i64 npairs,
                                                   i64 kk, i64 j) {
  for (i64 t = gtid(); t < npairs; t += gstride()) {
    const i64 lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
    const i64 hi = lo + j;
    if (hi >= len) continue;
    const bool up = (hi & kk) == 0;
    const float a = key[hi], b = key[lo];
    if ((a < b) ^ up) {
      key[hi] = b;
      key[lo] = a;
      const i32 tt = idx[hi];
      idx[hi] = idx[lo];
      idx[lo] = tt;
    }
  }
}

// ----------------------------------------------------------------- selection --
// gpu_part.cu:373-380 under SEQ: move at sorted position i is selected iff its
// rank among earlier positions with the same destination is below
// avail[dst].  Padding that reached the prefix is not selected.
constexpr int SEL_BLK = 1024;

__global__ void k_sel_dst(const i32* sk_idx, const i32* mv_dst, i64 M, i32* sdst) {
  for (i64 i = gtid(); i < M; i += gstride()) {
    const i32 id = sk_idx[i];
    sdst[i] = id >= 0 ? mv_dst[id] : -1;
  }
}

__global__ __launch_bounds__(SEL_BLK) void k_sel_hist(const i32* sdst, i64 M, i32 k, i32* H) {
  __shared__ i32 h[MAXK_LDS];
  for (int q = threadIdx.x; q < k; q += SEL_BLK) h[q] = 0;
  __syncthreads();
  const i64 i = (i64)blockIdx.x * SEL_BLK + threadIdx.x;
  if (i < M) {
    const i32 d = sdst[i];
    if (d >= 0 && d < k) atomicAdd(&h[d], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < k; q += SEL_BLK) H[(i64)blockIdx.x * k + q] = h[q];
}

__global__ void k_sel_scan(i32* H, i64 nb, i32 k) {
  for (i64 q = gtid(); q < k; q += gstride()) {
    i32 run = 0;
    for (i64 b = 0; b < nb; b++) {
      const i32 v = H[b * k + q];
      H[b * k + q] = run;
      run += v;
    }
  }
}

__global__ __launch_bounds__(SEL_BLK) void k_sel_mark(const i32* sdst, i64 M, i32 k,
                                                      const i32* H, const i64* avail, u8* sel,
                                                      i32* S, i32* pads) {
  __shared__ i32 cntd[MAXK_LDS];
  __shared__ i32 ssel[SEL_BLK / WAVE];
  for (int q = threadIdx.x; q < k; q += SEL_BLK) cntd[q] = H[(i64)blockIdx.x * k + q];
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const i64 i = (i64)blockIdx.x * SEL_BLK + threadIdx.x;
  const i32 d = i < M ? sdst[i] : -1;
  const bool valid = d >= 0 && d < k;
  bool mine = false;
  for (int step = 0; step < SEL_BLK / WAVE; step++) {
    if (wv == step) {
      u64 active = __ballot(valid);
      while (active) {
        const int leader = __ffsll((unsigned long long)active) - 1;
        const i32 dl = __shfl(d, leader);
        const u64 same = __ballot(valid && d == dl);
        const i32 basev = cntd[dl];
        __builtin_amdgcn_wave_barrier();
        if (valid && d == dl) {
          const i64 r = (i64)basev + __popcll(same & lanes_below(lane));
          mine = r < avail[dl];
        }
        if (lane == leader) cntd[dl] = basev + __popcll(same);
        __builtin_amdgcn_wave_barrier();
        active &= ~same;
      }
    }
    __syncthreads();
  }
  if (i < M) sel[i] = mine ? 1 : 0;
  const u64 pm = __ballot(i < M && !valid);
  if (lane == 0 && pm) atomicAdd(pads, (i32)__popcll(pm));
  const u64 bm = __ballot(mine);
  if (lane == 0) ssel[wv] = __popcll(bm);
  __syncthreads();
  if (threadIdx.x == 0) {
    i32 s = 0;
    for (int w = 0; w < SEL_BLK / WAVE; w++) s += ssel[w];
    if (s) atomicAdd(S, s);
  }
}

// gpu_part.cu:383-397.
__global__ void k_commit(const i32* sk_idx, const u8* sel, i64 M, const i32* mv_v,
                         const i32* mv_src, const i32* mv_dst, i32* part, i32* pc, i32 k) {
  for (i64 i = gtid(); i < M; i += gstride()) {
    if (!sel[i]) continue;
    const i32 id = sk_idx[i];
    if (id < 0) continue;
    const i32 src = mv_src[id], dst = mv_dst[id];
    if (src < k && dst < k && src >= 0 && dst >= 0) {
      const i32 v = mv_v[id];
      const i32 old = part[v];
      part[v] = dst;
      if (old == src) {
        atomicAdd(&pc[dst], 1);
        atomicAdd(&pc[src], -1);
      } else {
        part[v] = old;
      }
    }
  }
}

// -------------------------------------------------------------------- cut --
// gpu_part.cu:400-422: sum of w over arcs whose ends lie in different parts.
// Summed in uint64 (the int32 wrap of the reference is its low 32 bits).
__device__ __forceinline__ void block_add_u64(u64 v, u64* acc) {
  __shared__ u64 sw[16];
  v = wave_sum(v);
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  if (lane == 0) sw[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sw[w];
    if (s) atomicAdd((unsigned long long*)acc, (unsigned long long)s);
  }
}

__global__ __launch_bounds__(256) void k_cut_thread(DGraph g, i64 nedges, const i32* part,
                                                    i32 k, u64* acc, i64 heavy_deg) {
  u64 s = 0;
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > heavy_deg) continue;
    const i32 pv = part[v];
    if (pv < 0 || pv >= k) continue;
    for (i64 i = s0; i < s1; i++) {
      const i32 u = g.adj[i];
      if (u < g.n) {
        const i32 pu = part[u];
        if (pu != pv && pu >= 0 && pu < k) s += (u64)(i64)g.w[i];
      }
    }
  }
  block_add_u64(s, acc);
}

__global__ __launch_bounds__(256) void k_cut_heavy(DGraph g, i64 nedges, const i32* part,
                                                   i32 k, u64* acc, const i32* list, i64 cnt) {
  u64 s = 0;
  const int lane = lane_id();
  for (i64 e = gwave(); e < cnt; e += nwaves()) {
    const i32 v = list[e];
    const i32 pv = part[v];
    if (pv < 0 || pv >= k) continue;
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < g.n) {
        const i32 pu = part[u];
        if (pu != pv && pu >= 0 && pu < k) s += (u64)(i64)g.w[i];
      }
    }
  }
  block_add_u64(s, acc);
}

// Arc -> source vertex, built once per level for the arc-parallel kernels.
// Only arcs below the effective end of the last row (its range ends at the
// caller's nedges, gpu_part.cu:408) are sources of work; arcs past it keep -1.
__global__ __launch_bounds__(256) void k_arc_source(DGraph g, i32* asrc) {
  const int lane = lane_id();
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) asrc[i] = (i32)v;
  }
}

// gpu_part.cu:400-422 as an arc-parallel stream over [0, e_end): every arc of
// a row in the reference's range is visited once; sum in uint64 (its low 32
// bits are the reference's int32 wrap).
template <bool UNIT>
__global__ __launch_bounds__(256) void k_cut_arcs(const i32* asrc, const i32* adj, const i32* w,
                                                  i64 n, i64 e_end, const i32* part, i32 k,
                                                  u64* acc) {
  u64 s = 0;
  for (i64 i = gtid(); i < e_end; i += gstride()) {
    const i32 v = asrc[i];
    const i32 u = adj[i];
    if (u < n) {
      const i32 pv = part[v], pu = part[u];
      if (pv != pu && pv >= 0 && pv < k && pu >= 0 && pu < k) s += UNIT ? 1ull : (u64)(i64)w[i];
    }
  }
  block_add_u64(s, acc);
}

}  // namespace

// ================================================================ launchers ==
i64 scan_scratch_elems(i64 n) {
  i64 tot = 0;
  while (n > 1) {
    n = (n + SCAN_TILE - 1) / SCAN_TILE;
    tot += 2 * n;
  }
  return tot + 2;
}

void launch_fill_i32(i32* p, i64 n, i32 v, hipStream_t s) {
  if (n > 0) k_fill<i32><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(p, n, v);
}
void launch_fill_i64(i64* p, i64 n, i64 v, hipStream_t s) {
  if (n > 0) k_fill<i64><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(p, n, v);
}
void launch_fill_u8(u8* p, i64 n, u8 v, hipStream_t s) {
  if (n > 0) k_fill<u8><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(p, n, v);
}
void launch_fill_u64(u64* p, i64 n, u64 v, hipStream_t s) {
  if (n > 0) k_fill<u64><<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(p, n, v);
}

void launch_scan_excl(const i64* in, i64* out, i64 n, i64* total, ScanScratch sc,
                      hipStream_t s) {
  if (n <= 0) {
    k_fill<i64><<<1, 1, 0, s>>>(total, 1, 0);
    return;
  }
  const i64 nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  i64* bsum = sc.buf;
  i64* boff = sc.buf + nb;
  k_scan_tile<<<(unsigned)nb, SCAN_BS, 0, s>>>(in, out, n, bsum);
  if (nb > 1) {
    ScanScratch sub{sc.buf + 2 * nb, sc.cap - 2 * nb};
    launch_scan_excl(bsum, boff, nb, total, sub, s);
    k_scan_add<<<(unsigned)nb, SCAN_BS, 0, s>>>(out, n, boff);
  } else {
    k_copy1<<<1, 1, 0, s>>>(bsum, total);
  }
}

void launch_level_census(DGraph g, i64 nedges, i64 heavy_deg, i32* heavy_list,
                         i32* heavy_count, i32* nonunit, hipStream_t s) {
  k_level_census<<<grid_for(g.n > g.m ? g.n : g.m, 256, 8192), 256, 0, s>>>(
      g, nedges, heavy_deg, heavy_list, heavy_count, nonunit);
}

void launch_match_init(i64 n, i32* mate, u8* dec, i32* pend, u64* R, i64* lo,
                       hipStream_t s) {
  k_match_init<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(n, mate, dec, pend, R);
  k_fill<i64><<<1, 1, 0, s>>>(lo, 1, 0);
}

void launch_match_round(DGraph g, i32* mate, u8* dec, i32* pend, u64* R, i32* prop, i64* lo,
                        i64 W, u32 round, hipStream_t s) {
  k_match_reserve<<<grid_for(W, MATCH_BS, 4096), MATCH_BS, 0, s>>>(g, mate, dec, pend, R, prop,
                                                                   lo, W, round);
  k_match_decide<<<grid_for(W, MATCH_BS, 4096), MATCH_BS, 0, s>>>(R, prop, pend, mate, dec, lo,
                                                                  W, g.n, round);
}

void launch_match_sym_init(i64 n, i32* mate, u8* dec, i32* blkA, i32* blkC, i32* prop, i64* lo,
                           hipStream_t s) {
  k_match_sym_init<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(n, mate, dec, blkA, blkC, prop);
  k_fill<i64><<<1, 3, 0, s>>>(lo, 3, 0);
}

void launch_match_sym_round(DGraph g, i32* mate, u8* dec, i32* blkA, i32* blkC, i32* prop,
                            i32* decision, i64* lo, i64 W, i32* req, i32* nreq, u32 round,
                            hipStream_t s) {
  i32* cur = nreq + (round & 1);
  i32* next = nreq + ((round + 1) & 1);
  const int par = (int)(round & 1);
  k_match_sym_decide<<<grid_for(W * (WAVE / SYM_VPW), MATCH_BS, 4096), MATCH_BS, 0, s>>>(
      g, mate, dec, blkA, blkC, prop, decision, lo, W, req, cur, next, par);
  k_match_sym_apply<<<grid_for(W, MATCH_BS, 4096), MATCH_BS, 0, s>>>(decision, mate, dec, lo, W,
                                                                     g.n, par);
  k_match_huge<<<256, 1024, 0, s>>>(g, mate, dec, blkA, blkC, prop, req, cur);
}

void launch_sym_check(DGraph g, u64* table, i64 cap, i32* missing, hipStream_t s) {
  k_fill<u64><<<grid_for(cap, 256, 1u << 16), 256, 0, s>>>(table, cap, ~0ull);
  k_sym_insert<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, cap);
  k_sym_check<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, cap, missing);
}

void launch_relabel(const i32* mate, i64 n, i64* flags, i64* pos, i64* total, i32* new_ids,
                    i32* tb, ScanScratch sc, hipStream_t s) {
  k_root_flags<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(mate, n, flags);
  launch_scan_excl(flags, pos, n, total, sc, s);
  k_relabel<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(mate, pos, n, new_ids, tb);
}

void launch_contract_ref(DGraph g, const i32* mate, const i32* new_ids, i64 n_c, i32* sadj,
                         i32* sw, i64* cnt, hipStream_t s) {
  k_contract_ref<<<grid_for(n_c, 1, 1u << 18), WAVE, 0, s>>>(g, mate, new_ids, n_c, sadj, sw,
                                                             cnt);
}
void launch_compact_rows(const i32* sadj, const i32* sw, const i64* ox, i64 n_c, i32* oa,
                         i32* ow, hipStream_t s) {
  k_compact_rows<<<grid_for(n_c, 4, 1u << 16), 256, 0, s>>>(sadj, sw, ox, n_c, oa, ow);
}

void launch_rep_sizes(const i64* xadj, const i32* tb, i64 n_c, i64* dsz, i64* csz,
                      hipStream_t s) {
  k_rep_sizes<<<grid_for(n_c, 256, 1u << 16), 256, 0, s>>>(xadj, tb, n_c, dsz, csz);
}
void launch_rep_insert(DGraph g, const i32* tb, const i32* new_ids, i64 n_c, const i64* toff,
                       const i64* csz, i32* tkey, i32* tpos, i32* tw, hipStream_t s) {
  k_rep_insert<<<grid_for(n_c, 4, 1u << 16), 256, 0, s>>>(g, tb, new_ids, n_c, toff, csz, tkey,
                                                          tpos, tw);
}
void launch_rep_mark(DGraph g, const i32* tb, const i32* new_ids, i64 n_c, const i64* toff,
                     const i64* csz, const i64* eoff, const i32* tkey, const i32* tpos,
                     i32* eslot, i64* cnt, hipStream_t s) {
  k_rep_mark<<<grid_for(n_c, 4, 1u << 16), 256, 0, s>>>(g, tb, new_ids, n_c, toff, csz, eoff,
                                                        tkey, tpos, eslot, cnt);
}
void launch_rep_emit(const i64* eoff, const i64* dsz, const i32* eslot, const i64* toff,
                     const i32* tkey, const i32* tw, const i64* ox, i64 n_c, i32* oa, i32* ow,
                     hipStream_t s) {
  k_rep_emit<<<grid_for(n_c, 4, 1u << 16), 256, 0, s>>>(eoff, dsz, eslot, toff, tkey, tw, ox,
                                                        n_c, oa, ow);
}
void launch_pairs(DGraph g, u64* table, i64 cap, u64* count, hipStream_t s) {
  k_pairs<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, cap, count);
}

void launch_project(const i32* tb, const i32* part_c, i64 n_c, i32* part_f, i64 n_f, i32 k,
                    hipStream_t s) {
  k_project<<<grid_for(n_c, 256, 1u << 16), 256, 0, s>>>(tb, part_c, n_c, part_f, n_f, k);
}
void launch_sanitize_hist(i32* part, i64 n, i32 k, i32* pc, hipStream_t s) {
  k_sanitize_hist<<<grid_for(n, 256 * 16, 2048), 256, 0, s>>>(part, n, k, pc);
}
void launch_block_init(i32* part, i64 n, i32 k, hipStream_t s) {
  k_block_init<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(part, n, k);
}

void launch_bg_table(const i32* pc, i32 k, i64 n, int phase, float* bg, u8* src_big,
                     hipStream_t s) {
  k_bg_table<<<grid_for(k, 256, 64), 256, 0, s>>>(pc, k, n, phase, bg, src_big);
}

i64 gain_gscratch_elems(i32 k) { return k <= MAXK_LDS ? 0 : (i64)65536 * k; }

void launch_bg_order(const float* bg, i32 k, i32* ord, i32* run_end, hipStream_t s) {
  k_bg_order<<<1, 1024, 0, s>>>(bg, k, ord, run_end);
}

void launch_split_heavy(DGraph g, i64 nedges, const i32* heavy, i64 nheavy, i32* medium,
                        i32* nmedium, i32* huge, i32* nhuge, hipStream_t s) {
  if (nheavy > 0)
    k_split_heavy<<<grid_for(nheavy, 256, 4096), 256, 0, s>>>(g, nedges, heavy, nheavy, medium,
                                                              nmedium, huge, nhuge);
}

void launch_gain(DGraph g, i64 nedges, const i32* part, i32 k, const float* bg,
                 const u8* src_big, float* gain, i32* target, u8* cand, int phase, int unit,
                 i32* gscratch, const i32* medium, i64 nmedium, const i32* huge, i64 nhuge,
                 hipStream_t s) {
  if (k <= 512) {
    // light rows: 8 lanes per vertex, 32 tallies of k ints per block
    const size_t shm8 = (size_t)(2 * k + 32 * k) * 4;
    const unsigned gr8 = grid_for(g.n, 32, 1u << 16);
    if (unit)
      k_gain_group<8, true><<<gr8, 256, shm8, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                                    cand, phase, nullptr, 0, GAIN_LIGHT);
    else
      k_gain_group<8, false><<<gr8, 256, shm8, s>>>(g, nedges, part, k, bg, src_big, gain,
                                                     target, cand, phase, nullptr, 0, GAIN_LIGHT);
    if (nmedium > 0) {
      const size_t shm64 = (size_t)(2 * k + 4 * k) * 4;
      const unsigned gr64 = grid_for(nmedium, 4, 1u << 16);
      if (unit)
        k_gain_group<64, true><<<gr64, 256, shm64, s>>>(g, nedges, part, k, bg, src_big, gain,
                                                         target, cand, phase, medium, nmedium, 0);
      else
        k_gain_group<64, false><<<gr64, 256, shm64, s>>>(g, nedges, part, k, bg, src_big, gain,
                                                          target, cand, phase, medium, nmedium, 0);
    }
    if (nhuge > 0) {
      const size_t shmh = (size_t)(3 * k) * 4;
      const unsigned grh = grid_for(nhuge, 1, 8192);
      if (unit)
        k_gain_huge<true><<<grh, HUGE_BS, shmh, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                                     cand, phase, huge, nhuge);
      else
        k_gain_huge<false><<<grh, HUGE_BS, shmh, s>>>(g, nedges, part, k, bg, src_big, gain,
                                                      target, cand, phase, huge, nhuge);
    }
    return;
  }
  // large k: one wave per vertex for every vertex
  const bool lds = k <= MAXK_LDS;
  const size_t shm = lds ? (size_t)k * 12 : 0;
  const unsigned gr = 65536;
  if (lds) {
    if (unit)
      k_gain<true, true><<<gr, WAVE, shm, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                               cand, phase, gscratch, nullptr, g.n, nullptr,
                                               nullptr);
    else
      k_gain<false, true><<<gr, WAVE, shm, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                                cand, phase, gscratch, nullptr, g.n, nullptr,
                                                nullptr);
  } else {
    if (unit)
      k_gain<true, false><<<gr, WAVE, 0, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                              cand, phase, gscratch, nullptr, g.n, nullptr,
                                              nullptr);
    else
      k_gain<false, false><<<gr, WAVE, 0, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                               cand, phase, gscratch, nullptr, g.n, nullptr,
                                               nullptr);
  }
}

void launch_token(DGraph g, i64 nedges, const float* gain, const u8* cand, const i32* part,
                  const i32* target, i32 k, i64* flag, const i32* heavy_list, i64 heavy_count,
                  i64 heavy_deg, hipStream_t s) {
  k_token_thread<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, nedges, gain, cand, part,
                                                              target, k, flag, heavy_deg);
  if (heavy_count > 0)
    k_token_heavy<<<grid_for(heavy_count, 4, 8192), 256, 0, s>>>(g, nedges, gain, cand, part,
                                                                 target, k, flag, heavy_list,
                                                                 heavy_count);
}

void launch_scatter_moves(const i64* flag, const i64* pos, const i32* part, const i32* target,
                          const float* gain, i64 n, i32 k, i32* mv_v, i32* mv_src, i32* mv_dst,
                          float* mv_gain, i32* outc, hipStream_t s) {
  k_scatter_moves<<<grid_for(n, 256 * 8, 4096), 256, 0, s>>>(flag, pos, part, target, gain, n,
                                                             k, mv_v, mv_src, mv_dst, mv_gain,
                                                             outc);
}

void launch_capacity(const i32* pc, const i32* outc, i32 k, i64 n, int phase, i64* avail,
                     hipStream_t s) {
  k_capacity<<<grid_for(k, 256, 64), 256, 0, s>>>(pc, outc, k, n, phase, avail);
}

void launch_sort_moves(const float* mv_gain, i64 M, i64 len, float* key, i32* idx,
                       hipStream_t s) {
  k_sort_init<<<grid_for(len, 256, 1u << 16), 256, 0, s>>>(mv_gain, M, len, key, idx);
  if (len < 2) return;
  const unsigned tiles = (unsigned)((len + BT_ELEMS - 1) / BT_ELEMS);
  k_bt_local<<<tiles, BT_THREADS, 0, s>>>(key, idx, len, 0);
  i64 p2 = 1;
  while (p2 < len) p2 <<= 1;
  const i64 npairs = p2 >> 1;
  for (i64 kk = 2 * BT_ELEMS; kk <= len; kk <<= 1) {
    for (i64 j = kk >> 1; j >= BT_ELEMS; j >>= 1)
      k_bt_global<<<grid_for(npairs, 256, 1u << 16), 256, 0, s>>>(key, idx, len, npairs, kk, j);
    k_bt_local<<<tiles, BT_THREADS, 0, s>>>(key, idx, len, kk);
  }
}

i64 select_scratch_elems(i64 M, i32 k) { return ((M + SEL_BLK - 1) / SEL_BLK + 1) * (i64)k; }

void launch_select(const i32* sk_idx, const i32* mv_dst, i64 M, i32 k, const i64* avail,
                   i32* sdst, i32* H, u8* sel, i32* S, i32* pads, hipStream_t s) {
  const i64 nb = (M + SEL_BLK - 1) / SEL_BLK;
  k_sel_dst<<<grid_for(M, 256, 1u << 16), 256, 0, s>>>(sk_idx, mv_dst, M, sdst);
  k_sel_hist<<<(unsigned)nb, SEL_BLK, 0, s>>>(sdst, M, k, H);
  k_sel_scan<<<grid_for(k, 64, 1024), 64, 0, s>>>(H, nb, k);
  k_sel_mark<<<(unsigned)nb, SEL_BLK, 0, s>>>(sdst, M, k, H, avail, sel, S, pads);
}

void launch_commit(const i32* sk_idx, const u8* sel, i64 M, const i32* mv_v, const i32* mv_src,
                   const i32* mv_dst, i32* part, i32* pc, i32 k, hipStream_t s) {
  k_commit<<<grid_for(M, 256, 1u << 16), 256, 0, s>>>(sk_idx, sel, M, mv_v, mv_src, mv_dst,
                                                      part, pc, k);
}

void launch_arc_source(DGraph g, i32* asrc, hipStream_t s) {
  k_arc_source<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, asrc);
}

void launch_cut_arcs(DGraph g, i64 e_end, const i32* asrc, const i32* part, i32 k, int unit,
                     u64* acc, hipStream_t s) {
  const unsigned gr = grid_for(e_end, 256 * 8, 8192);
  if (unit)
    k_cut_arcs<true><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part, k, acc);
  else
    k_cut_arcs<false><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part, k, acc);
}

void launch_cut(DGraph g, i64 nedges, const i32* part, i32 k, u64* acc, const i32* heavy_list,
                i64 heavy_count, i64 heavy_deg, hipStream_t s) {
  k_cut_thread<<<grid_for(g.n, 256, 8192), 256, 0, s>>>(g, nedges, part, k, acc, heavy_deg);
  if (heavy_count > 0)
    k_cut_heavy<<<grid_for(heavy_count, 4, 8192), 256, 0, s>>>(g, nedges, part, k, acc,
                                                               heavy_list, heavy_count);
}

}  // namespace tgp
