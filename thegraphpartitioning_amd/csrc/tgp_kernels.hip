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

#include <algorithm>
#include <climits>

namespace tgp {

namespace {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }
// row stride of a kept tally array: k rounded up to 4 ints (16-byte rows)
__host__ __device__ __forceinline__ i64 tally_stride(i32 k) { return (i64)((k + 3) & ~3); }
// byte rows of the kept tallies: k rounded up to 16
__host__ __device__ __forceinline__ i64 tally8_stride_d(i32 k) { return (i64)((k + 15) & ~15); }
// short (16-bit) rows of the kept tallies: k rounded up to 8
__host__ __device__ __forceinline__ i64 tally16_stride_d(i32 k) { return (i64)((k + 7) & ~7); }
// int row of v, or -1 for a byte row
__device__ __forceinline__ i64 tally_int_row(const TallyRef& T, i64 v) {
  return T.all_int ? v : (i64)T.bix[v];
}
__device__ __forceinline__ void tally_put(const TallyRef& T, i64 v, i32 k, i32 q, i32 val) {
  const i64 b = tally_int_row(T, v);
  if (b >= 0) T.t32[(i64)b * tally_stride(k) + q] = val;
  else if (T.t16) T.t16[v * tally16_stride_d(k) + q] = (u16)val;
  else T.t8[v * tally8_stride_d(k) + q] = (u8)val;
}
// A move of weight w from part src to part dst seen from neighbour u:
// T[u][src] -= w, T[u][dst] += w.  A byte row's adds go to the 32-bit words
// holding the bytes: every byte stays in [0, 255], so nothing carries into or
// borrows from its neighbours ((u32)d << sh is d * 2^sh modulo 2^32 for
// either sign), and two bytes of one word take one add of both terms.  The
// row class comes from the bigmask bit (an L2-resident bit array) before any
// int-row index is read.
__device__ __forceinline__ void tally_move(const TallyRef& T, i64 u, i32 k, i32 src, i32 dst,
                                          i32 w) {
  const bool big = T.all_int || ((T.bigmask[u >> 5] >> (u & 31)) & 1u);
  if (big) {
    i32* row = T.t32 + (T.all_int ? u : (i64)T.bix[u]) * tally_stride(k);
    atomicAdd(&row[src], -w);
    atomicAdd(&row[dst], w);
  } else if (T.t16) {  // short row: the same on 16-bit halves of a word
    u16* row = T.t16 + u * tally16_stride_d(k);
    const u32 ds = (u32)(-w) << (16 * (u32)(src & 1)), dd = (u32)w << (16 * (u32)(dst & 1));
    if ((src >> 1) == (dst >> 1)) {
      atomicAdd((u32*)(row + (src & ~1)), ds + dd);
    } else {
      atomicAdd((u32*)(row + (src & ~1)), ds);
      atomicAdd((u32*)(row + (dst & ~1)), dd);
    }
  } else {
    u8* row = T.t8 + u * tally8_stride_d(k);
    const u32 ds = (u32)(-w) << (8 * (u32)(src & 3)), dd = (u32)w << (8 * (u32)(dst & 3));
    if ((src >> 2) == (dst >> 2)) {
      atomicAdd((u32*)(row + (src & ~3)), ds + dd);
    } else {
      atomicAdd((u32*)(row + (src & ~3)), ds);
      atomicAdd((u32*)(row + (dst & ~3)), dd);
    }
  }
}
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

// Block-wide sum of one u64 per thread, added by thread 0 to one of the
// ACC_SLOTS slots of `acc` (chosen by block index, one 128-byte line each) so
// that thousands of blocks do not serialise on a single address; k_acc_finish
// folds the slots.  Every thread of the block must call it.
__device__ __forceinline__ void block_add_u64(u64 v, u64* acc) {
  __shared__ u64 sw[16];
  v = wave_sum(v);
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  if (lane == 0) sw[wv] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sw[w];
    if (s)
      atomicAdd((unsigned long long*)&acc[(blockIdx.x & (ACC_SLOTS - 1)) * ACC_STRIDE],
                (unsigned long long)s);
  }
}

// list[(*cnt)++] = x for every active lane with pred, in lane order within the
// wave, with one atomic per wave (callable from divergent code).
__device__ __forceinline__ void wave_append(bool pred, i32 x, i32* list, i32* cnt) {
  const u64 m = __ballot(pred);
  if (!m) return;
  const int lane = lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  i32 base = 0;
  if (lane == leader) base = atomicAdd(cnt, (i32)__popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1))] = x;
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
                                                      i32* nonunit, u64* maxabs) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    wave_append(s1 - s0 > heavy_deg, (i32)v, heavy_list, heavy_count);
  }
  bool nu = false;
  u64 mx = 0;
  for (i64 i = gtid(); i < g.m; i += gstride()) {
    const i32 w = g.w[i];
    nu |= w != 1;
    const u64 a = w < 0 ? (u64)(-(i64)w) : (u64)w;
    mx = a > mx ? a : mx;
  }
  if (nu) atomicOr(nonunit, 1);
  mx = wave_max_u64(mx);
  if (lane_id() == 0 && mx) atomicMax((unsigned long long*)maxabs, (unsigned long long)mx);
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
constexpr int SCAN_U = 32;
constexpr int TU = 16;  // loads per lane per step of the proposal scans (fewer registers)

// Device-scope (coherent) accesses to the matching state, for kernels in
// which other waves update it while they run.
__device__ __forceinline__ u8 ld_dec(const u8* dec, i64 i) {
  return __hip_atomic_load(const_cast<u8*>(dec) + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ i32 ld_mate(const i32* mate, i64 i) {
  return __hip_atomic_load(const_cast<i32*>(mate) + i, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dec(u8* dec, i64 i) {
  __hip_atomic_store(dec + i, (u8)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_mate(i32* mate, i64 i, i32 v) {
  __hip_atomic_store(mate + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool COH>
__device__ __forceinline__ u8 get_dec(const u8* dec, i64 i) {
  return COH ? ld_dec(dec, i) : dec[i];
}
template <bool COH>
__device__ __forceinline__ i32 get_mate(const i32* mate, i64 i) {
  return COH ? ld_mate(mate, i) : mate[i];
}

// largest undecided neighbour u of x with u < lim, u != x (BLK_NONE if none)
template <bool COH = false>
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
    for (int j = 0; j < SCAN_U; j++)
      d[j] = (u[j] < lim && u[j] != x) ? get_dec<COH>(dec, u[j]) : (u8)1;
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

template <bool COH = false>
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
      mt[j] = (u[j] >= 0 && u[j] < g.n && u[j] != t && w[j] >= 0) ? get_mate<COH>(mate, u[j]) : 0;
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

// Chained matching for symmetric levels: the decision rule of the rounds
// above (conditions A, B, C with cached blockers and proposals) without
// rounds.  The state is one word per vertex, st[x] = -1 (undecided), -2
// (decided, unmatched) or the partner, and changes once.  Waves take groups
// of CH_VPW consecutive vertices in ascending order from a ticket counter,
// keep each vertex's blocker and proposal in registers, and re-examine the
// group until every vertex is decided, publishing a commit t-u at once as
// st[t] = u, st[u] = t.  No ordering between the two stores is needed,
// because whoever must know that u is taken observes it in a word it has to
// read anyway: a vertex that waits for its lower neighbours (A) sees a
// neighbour y with st[y] == itself when y took it, and a vertex that waits for
// the lower neighbours of its proposal u (C) sees st[y] == u when y took u.
// Every condition is monotone in the decided set, so a decision taken on a
// later view than another wave's is still the sequential one; the only
// requirement on the memory system is per-word coherence (device-scope
// relaxed atomics).  Two vertices of one wave never commit on each other: the
// larger waits for the smaller by A or C.  The minimum undecided vertex always
// passes A-C and groups are handed out in ascending order, so the group that
// holds it makes progress.  Every wave also stops when the launch's time
// budget is spent; the host then launches again from the first undecided
// vertex, with the caches written back.
constexpr i32 ST_UNDEC = -1;
constexpr i32 ST_EXHAUST = -2;
constexpr i32 BLK_TAKEN = -3;  // scan result: the scanned word names the asker
constexpr i32 PROP_NONE = -5;  // proposal: no free neighbour (final)

__device__ __forceinline__ i32 ld_st(const i32* st, i64 i) { return ld_mate(st, i); }

// Among neighbours v of x with v < lim, v != x, at row positions from `from`
// on, one trip of WAVE * SCAN_U entries at a time: BLK_TAKEN if a decided v of
// the trip has st[v] == me, else the largest undecided v of the first trip
// that holds one (`from` is set to that trip's start), else BLK_NONE (`from`
// set to the row's end).  A decided word never changes, so entries before
// `from` stay decided and not equal to me: a later scan of the same question
// (same x, lim, me) resumes there, and every long row is read about once
// per question instead of once per blocker.
// The entries scanned are adj[s0, s1): the row of x, or its lower-neighbour
// list (ChainLow), which holds every entry of the row below x.
__device__ __forceinline__ i32 wave_scan_pending(const i32* adj, i64 s0, i64 s1, i32 x, i32 lim,
                                                 i32 me, const i32* st, int lane, i32& from,
                                                 bool sorted = false) {
  for (i64 base = s0 + from; base < s1; base += WAVE * SCAN_U) {
    i32 v[SCAN_U];
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) {
      const i64 i = base + j * WAVE + lane;
      v[j] = i < s1 ? adj[i] : INT_MAX;
    }
    i32 sv[SCAN_U];
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) sv[j] = (v[j] < lim && v[j] != x) ? ld_st(st, v[j]) : -4;
    i32 best = BLK_NONE;
    bool taken = false;
#pragma unroll
    for (int j = 0; j < SCAN_U; j++) {
      if (sv[j] == ST_UNDEC) best = v[j] > best ? v[j] : best;
      taken |= sv[j] == me;
    }
    if (__ballot(taken)) return BLK_TAKEN;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const i32 t2 = __shfl_xor(best, o);
      best = t2 > best ? t2 : best;
    }
    if (best != BLK_NONE) {
      from = (i32)(base - s0);
      return best;
    }
    // ascending entries: once a trip reaches lim, no later entry is below it
    if (sorted && __ballot(v[SCAN_U - 1] >= lim)) break;
  }
  from = (i32)(s1 - s0);
  return BLK_NONE;
}

// Unit-weight levels: the heaviest free neighbour of t (first position on
// ties) is its first free entry, v != t, 0 <= v < n, st[v] < 0.  Entries
// before `from` are not free and stay so (a taken vertex stays taken), so the
// scan resumes there; returns the neighbour (-1 if none) with its state in
// sval, and leaves `from` at its position.
__device__ __forceinline__ i32 wave_first_free(const DGraph& g, i32 t, const i32* st, int lane,
                                               i32& from, i32& sval) {
  const i64 s0 = g.xadj[t], s1 = g.xadj[t + 1];
  for (i64 base = s0 + from; base < s1; base += WAVE * TU) {
    i32 v[TU], sv[TU];
#pragma unroll
    for (int j = 0; j < TU; j++) {
      const i64 i = base + j * WAVE + lane;
      v[j] = i < s1 ? g.adj[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < TU; j++)
      sv[j] = (v[j] >= 0 && v[j] < g.n && v[j] != t) ? ld_st(st, v[j]) : 0;
    i32 key = INT_MAX, myv = -1, mys = 0;
#pragma unroll
    for (int j = TU - 1; j >= 0; j--) {
      if (sv[j] < 0) {
        key = j * WAVE + lane;
        myv = v[j];
        mys = sv[j];
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const i32 k2 = __shfl_xor(key, o);
      key = k2 < key ? k2 : key;
    }
    if (key != INT_MAX) {
      const int src = key & (WAVE - 1);
      from = (i32)(base - s0) + key;
      sval = __shfl(mys, src);
      return __shfl(myv, src);
    }
  }
  from = (i32)(s1 - s0);
  sval = 0;
  return -1;
}

// The K best free (st < 0) neighbours of t by (weight, first position), as
// keys (w << 32 | ~position), best first: lane sgbase + r receives the r-th
// in `mykey` (0 where there are fewer), and `nvalid` (uniform) says how many
// of them are exact.  Every other free neighbour ranks below the valid ones,
// and a taken vertex stays taken, so the best free neighbour at any later
// time is the first still-free valid key (when one is left).  Each lane keeps
// its own 4 best; the wave merges them one key per round, and stops after a
// round that empties a lane which saw more than 4 (its fifth might be next).
template <int K>
__device__ __forceinline__ void wave_best_free_topk(const DGraph& g, i32 t, const i32* st,
                                                    int lane, int sgbase, u64& mykey,
                                                    int& nvalid) {
  const i64 s0 = g.xadj[t], s1 = g.xadj[t + 1];
  u64 k0 = 0, k1 = 0, k2 = 0, k3 = 0;
  int seen = 0;
  for (i64 base = s0; base < s1; base += WAVE * TU) {
    i32 v[TU], w[TU], sv[TU];
#pragma unroll
    for (int j = 0; j < TU; j++) {
      const i64 i = base + j * WAVE + lane;
      v[j] = i < s1 ? g.adj[i] : -1;
      w[j] = i < s1 ? g.w[i] : -1;
    }
#pragma unroll
    for (int j = 0; j < TU; j++)
      sv[j] = (v[j] >= 0 && v[j] < g.n && v[j] != t && w[j] >= 0) ? ld_st(st, v[j]) : 0;
#pragma unroll
    for (int j = 0; j < TU; j++) {
      if (sv[j] < 0) {
        seen++;
        const i64 i = base + j * WAVE + lane;
        u64 c = ((u64)(u32)w[j] << 32) | (u64)(0xFFFFFFFFu - (u32)(i - s0));
        if (c > k3) {
          k3 = c;
          if (k3 > k2) { c = k2; k2 = k3; k3 = c; }
          if (k2 > k1) { c = k1; k1 = k2; k2 = c; }
          if (k1 > k0) { c = k0; k0 = k1; k1 = c; }
        }
      }
    }
  }
  mykey = 0;
  nvalid = 0;
  int popped = 0;
#pragma unroll 1
  for (int r = 0; r < K; r++) {
    const u64 m = wave_max_u64(k0);
    if (m == 0) break;
    if (lane == sgbase + r) mykey = m;
    nvalid = r + 1;
    const bool win = k0 == m;  // positions are distinct: one lane holds m
    if (win) {
      k0 = k1;
      k1 = k2;
      k2 = k3;
      k3 = 0;
      popped++;
    }
    if (__ballot(win && popped == 4 && seen > 4)) break;
  }
}

// Best-entry lists of the long rows (ChainLow): rows longer than minlen get a
// list index.
__global__ void k_list_rows(DGraph g, i64 minlen, i32* lidx, i32* lrows, i32* nl) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const bool lng = g.xadj[v + 1] - g.xadj[v] > minlen;
    i32 idx = -1;
    if (lng) {
      idx = atomicAdd(nl, 1);
      lrows[idx] = (i32)v;
    }
    lidx[v] = idx;
  }
}

// One block per listed row.  Every thread keeps the LST_T best keys
// (w << 32 | ~position) of its entries i = t, t + 256, ... (sorted, in LDS
// afterwards); wave 0 then pops the largest kept key CHAIN_LIST_LEN times
// (lane l merges the lists of threads 4l .. 4l + 3).  The popped keys are
// the row's best keys exactly down to the largest key a thread dropped below
// (a thread that saw more than LST_T keeps its LST_T-th best, and every key it
// dropped is smaller), where the list stops.  The list holds the neighbour
// ids, best first.
constexpr int LST_T = 8;
constexpr int LST_BS = 256;
__global__ __launch_bounds__(LST_BS) void k_long_lists(DGraph g, const i32* lrows, const i32* nl,
                                                       i32* lists, i32* lcnt) {
  constexpr int NK = LST_BS * LST_T;
  constexpr int PER_LANE = LST_BS / WAVE;  // kept lists merged by one lane of wave 0
  __shared__ u64 sk[NK];
  __shared__ u64 sbound[LST_BS / WAVE];
  const int t = threadIdx.x, lane = lane_id(), wv = t >> 6;
  const i32 nrows = *nl;
  for (i32 li = blockIdx.x; li < nrows; li += gridDim.x) {
    const i32 x = lrows[li];
    const i64 s0 = g.xadj[x], s1 = g.xadj[x + 1];
    u64 k[LST_T];
#pragma unroll
    for (int j = 0; j < LST_T; j++) k[j] = 0;
    i64 seen = 0;
    constexpr int LST_U = 8;  // entries per thread in flight
    for (i64 base = s0 + t; base < s1; base += LST_BS * LST_U) {
      i32 v[LST_U], w[LST_U];
#pragma unroll
      for (int u = 0; u < LST_U; u++) {
        const i64 i = base + (i64)u * LST_BS;
        v[u] = i < s1 ? g.adj[i] : -1;
        w[u] = i < s1 ? g.w[i] : -1;
      }
#pragma unroll
      for (int u = 0; u < LST_U; u++) {
        if (v[u] < 0 || v[u] >= g.n || v[u] == x || w[u] < 0) continue;
        seen++;
        const i64 i = base + (i64)u * LST_BS;
        const u64 c = ((u64)(u32)w[u] << 32) | (u64)(0xFFFFFFFFu - (u32)(i - s0));
        if (c > k[LST_T - 1]) {
          k[LST_T - 1] = c;
#pragma unroll
          for (int j = LST_T - 1; j > 0; j--) {
            if (k[j] > k[j - 1]) {
              const u64 tmp = k[j];
              k[j] = k[j - 1];
              k[j - 1] = tmp;
            }
          }
        }
      }
    }
    u64 bnd = seen > LST_T ? k[LST_T - 1] : 0;
    bnd = wave_max_u64(bnd);
    if (lane == 0) sbound[wv] = bnd;
#pragma unroll
    for (int j = 0; j < LST_T; j++) sk[t * LST_T + j] = k[j];
    __syncthreads();
    if (wv == 0) {
      u64 B = 0;
#pragma unroll
      for (int w2 = 0; w2 < LST_BS / WAVE; w2++) B = sbound[w2] > B ? sbound[w2] : B;
      // lane's merge state: the next key of each of its PER_LANE kept lists
      int p[PER_LANE];
      u64 hd[PER_LANE];
#pragma unroll
      for (int q = 0; q < PER_LANE; q++) {
        p[q] = 0;
        hd[q] = sk[(lane * PER_LANE + q) * LST_T];
      }
      u64 out = 0;
      int cnt = 0;
      for (int r = 0; r < CHAIN_LIST_LEN; r++) {
        u64 mine = 0;
        int qb = 0;
#pragma unroll
        for (int q = 0; q < PER_LANE; q++) {
          if (hd[q] > mine) {
            mine = hd[q];
            qb = q;
          }
        }
        const u64 m = wave_max_u64(mine);
        if (m == 0 || m < B) break;
        if (lane == r) out = m;
        cnt = r + 1;
        if (mine == m) {  // keys are distinct: one lane holds m
#pragma unroll
          for (int q = 0; q < PER_LANE; q++) {
            if (q == qb) {
              p[q]++;
              hd[q] = p[q] < LST_T ? sk[(lane * PER_LANE + q) * LST_T + p[q]] : 0;
            }
          }
        }
      }
      lists[(i64)li * CHAIN_LIST_LEN + lane] =
          lane < cnt ? g.adj[s0 + (i64)(0xFFFFFFFFu - (u32)out)] : -1;
      if (lane == 0) lcnt[li] = cnt;
    }
    __syncthreads();
  }
}

// Sub-group layout: CH_SG lanes per vertex, CH_VPW vertices per wave.  A row
// of at most CH_ROW arcs is held in registers (CH_ROW / CH_SG entries per
// lane), so a rescan of it is one batch of state loads and all sub-groups of
// a wave rescan at once.  Longer rows are scanned by the whole wave, one row
// at a time.
template <int CH_SG>
__device__ __forceinline__ i32 sg_max_i32(i32 v) {
#pragma unroll
  for (int o = CH_SG / 2; o > 0; o >>= 1) {
    const i32 t = __shfl_xor(v, o, CH_SG);
    v = t > v ? t : v;
  }
  return v;
}
template <int CH_SG>
__device__ __forceinline__ bool sg_any(bool f) {
  i32 v = f;
#pragma unroll
  for (int o = CH_SG / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, CH_SG);
  return v != 0;
}

// The chained matching of one wave: groups from `ticket` starting at vertex
// lo, until the groups run out or the wall-clock budget is spent.
template <int CH_SG, int CH_PER, bool STATS = false>
__device__ __forceinline__ void chain_wave(const DGraph& g, const ChainLow& low, i32* st,
                                           i32* blkA, i32* blkC, i32* prop, const i64 lo,
                                           u32* ticket, i64 budget, u64* stats) {
  constexpr int CH_VPW = WAVE / CH_SG;
  constexpr int CH_ROW = CH_SG * CH_PER;
  constexpr u64 SG_MASK = CH_SG == 64 ? ~0ull : (1ull << (CH_SG % 64)) - 1;
  // stats (STATS): per-kind event counts, summed over the grid
  u64 c_iter = 0, c_sleep = 0, c_apS = 0, c_aL = 0, c_pL = 0, c_fetch = 0, c_cS = 0, c_cL = 0,
      c_commit = 0, c_arcsL = 0, c_arcsP = 0, c_arcsC = 0;
  const i64 n = g.n;
  const int lane = lane_id();
  const int sl = lane & (CH_SG - 1);
  const int sgbase = lane & ~(CH_SG - 1);
  const i64 t_start = wall_clock64();
  i64 base = -1;
  // per-vertex state, uniform within a sub-group
  bool own = false, tshort = false, ushort_ = false, uknown = false;
  // long rows: where the next A scan of t's row and C scan of u's row resume,
  // and (unit-weight levels) the next proposal scan of t's row; spos: the
  // first entry of t's best-entry list not known to be taken
  i32 apos = 0, cpos = 0, ppos = 0, spos = 0;
  i32 t = -1, a = BLK_NONE, u = -1, b = BLK_NONE;
  // long rows: lanes 0..pnv-1 of the sub-group hold the best free candidates
  // of the last full proposal scan (pci), ph = the entry now proposed (PK:
  // none)
  constexpr int PK = CH_SG < 16 ? CH_SG : 16;
  i32 pci = -1;
  int ph = PK, pnv = 0;
  i64 ts0 = 0;
  i32 tv[CH_PER], tw[CH_PER], uv[CH_PER];
  for (u32 iter = 0; iter < (1u << 26); iter++) {
    if (base < 0) {
      u32 tk = 0;
      if (lane == 0) tk = atomicAdd(ticket, 1u);
      tk = __shfl(tk, 0);
      base = lo + (i64)tk * CH_VPW;
      if (base >= n) break;
      const i64 e = base + (lane / CH_SG);
      own = e < n;
      t = own ? (i32)e : -1;
      a = own ? blkA[t] : BLK_NONE;
      u = own ? prop[t] : -1;
      b = own ? blkC[t] : BLK_NONE;
      uknown = false;
      cpos = 0;
      apos = 0;
      ppos = 0;
      spos = 0;
      pci = -1;
      ph = PK;
      pnv = 0;
      ts0 = own ? g.xadj[t] : 0;
      const i64 ts1 = own ? g.xadj[t + 1] : 0;
      tshort = own && ts1 - ts0 <= CH_ROW;
#pragma unroll
      for (int j = 0; j < CH_PER; j++) {
        const i64 i = ts0 + j * CH_SG + sl;
        tv[j] = (tshort && i < ts1) ? g.adj[i] : -1;
        tw[j] = (tshort && i < ts1) ? g.w[i] : -1;
      }
    }
    // one batch of coherent loads: lanes 0-3 of the sub-group read st of t, a, u, b
    i32 mine = ST_UNDEC;
    if (own) {
      const i32 x = sl == 0 ? t : sl == 1 ? a : sl == 2 ? u : sl == 3 ? b : -1;
      if (sl < 4 && x >= 0) mine = ld_st(st, x);
    }
    const i32 s_t = __shfl(mine, 0, CH_SG), s_a = __shfl(mine, 1, CH_SG);
    i32 s_u = __shfl(mine, 2, CH_SG);
    const i32 s_b = __shfl(mine, 3, CH_SG);
    own = own && s_t == ST_UNDEC;
    if (!__ballot(own)) {
      base = -1;
      continue;
    }
    if (wall_clock64() - t_start > budget) break;
    if constexpr (STATS) c_iter++;

    if (own && a >= 0 && s_a != ST_UNDEC) {
      if (s_a == t) own = false;  // a took t
      else a = BLK_UNKNOWN;
    }
    if (own && u >= 0 && s_u >= 0) u = BLK_UNKNOWN;  // proposal taken
    // A and proposal over a row held in registers: one batch of loads
    const bool needAP = own && (a == BLK_UNKNOWN || u == BLK_UNKNOWN);
    if (__ballot(needAP && tshort)) {
      if constexpr (STATS) c_apS++;
      i32 sv[CH_PER];
#pragma unroll
      for (int j = 0; j < CH_PER; j++)
        sv[j] = (needAP && tshort && tv[j] >= 0 && tv[j] < n && tv[j] != t) ? ld_st(st, tv[j])
                                                                          : -4;
      i32 bestA = BLK_NONE;
      bool tk = false;
      u64 key = 0;
      i32 kv = -1, ks = 0;
#pragma unroll
      for (int j = 0; j < CH_PER; j++) {
        if (sv[j] == -4) continue;
        if (tv[j] < t) {
          if (sv[j] == ST_UNDEC) bestA = tv[j] > bestA ? tv[j] : bestA;
          tk |= sv[j] == t;
        }
        if (sv[j] < 0 && tw[j] >= 0) {
          const u64 c = ((u64)(u32)tw[j] << 32) | (u64)(0xFFFFFFFFu - (u32)(j * CH_SG + sl));
          if (c > key) {
            key = c;
            kv = tv[j];
            ks = sv[j];
          }
        }
      }
      bestA = sg_max_i32<CH_SG>(bestA);
      tk = sg_any<CH_SG>(tk);
#pragma unroll
      for (int o = CH_SG / 2; o > 0; o >>= 1) {
        const u64 ok = __shfl_xor(key, o, CH_SG);
        const i32 ov = __shfl_xor(kv, o, CH_SG), os = __shfl_xor(ks, o, CH_SG);
        if (ok > key) {
          key = ok;
          kv = ov;
          ks = os;
        }
      }
      if (needAP && tshort) {
        if (a == BLK_UNKNOWN) {
          if (tk) own = false;
          else a = bestA;
        }
        if (u == BLK_UNKNOWN) {
          u = key ? kv : PROP_NONE;
          s_u = key ? ks : ST_UNDEC;
          b = BLK_UNKNOWN;
          uknown = false;
          cpos = 0;
        }
      }
    }
    // long rows: whole-wave scans, one sub-group at a time
    u64 need = __ballot(own && !tshort && sl == 0 && a == BLK_UNKNOWN);
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(t, l);
      i32 from = __shfl(apos, l);
      // every entry below x is on x's lower-neighbour list when it has one
      const i32* la = low.off ? low.adj : g.adj;
      const i64 l0 = low.off ? low.off[x] : g.xadj[x], l1 = low.off ? low.off[x + 1] : g.xadj[x + 1];
      if constexpr (STATS) c_aL++;
      if constexpr (STATS) c_arcsL += l1 - l0 - from;
      const i32 r = wave_scan_pending(la, l0, l1, x, x, x, st, lane, from, low.sorted);
      if ((lane & ~(CH_SG - 1)) == l) {
        if (r == BLK_TAKEN) own = false;
        else a = r;
        apos = from;
      }
      need &= need - 1;
    }
    // long rows: a taken proposal is replaced by the first still-free cached
    // candidate; only when none is left is the row scanned again
    {
      const bool tryc = own && !tshort && u == BLK_UNKNOWN && ph < pnv - 1;
      const bool mine = tryc && sl > ph && sl < pnv && pci >= 0;
      const i32 sv = mine ? ld_st(st, pci) : 0;
      const u64 fm = (__ballot(mine && sv < 0) >> (lane & ~(CH_SG - 1))) & SG_MASK;
      if (tryc) {
        if (fm) {
          const int r = __ffsll((unsigned long long)fm) - 1;
          u = __shfl(pci, r, CH_SG);
          s_u = __shfl(sv, r, CH_SG);
          b = BLK_UNKNOWN;
          uknown = false;
          cpos = 0;
          ph = r;
        } else {
          ph = PK;
        }
      }
    }
    need = __ballot(own && !tshort && sl == 0 && u == BLK_UNKNOWN);
    while (need && low.unitw) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(t, l);
      i32 from = __shfl(ppos, l), sr = 0;
      if constexpr (STATS) c_pL++;
      if constexpr (STATS) c_arcsP -= from;
      const i32 r0 = wave_first_free(g, x, st, lane, from, sr);
      if constexpr (STATS) c_arcsP += from;
      if ((lane & ~(CH_SG - 1)) == l) {
        ppos = from;
        u = r0 >= 0 ? r0 : PROP_NONE;
        s_u = r0 >= 0 ? sr : ST_UNDEC;
        b = BLK_UNKNOWN;
        uknown = false;
        cpos = 0;
      }
      need &= need - 1;
    }
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(t, l);
      if constexpr (STATS) c_pL++;
      if constexpr (STATS) c_arcsP += g.xadj[x + 1] - g.xadj[x];
      // x's best-entry list: its first entry still free (the entries before
      // spos are taken, and a taken vertex stays taken) is x's best free
      // neighbour; only when none is left is the row scanned
      const i32 li = low.lidx ? low.lidx[x] : -1;
      if (li >= 0) {
        const i32 from = __shfl(spos, l);
        const i32 nl = low.lcnt[li];
        const i32 e = (lane >= from && lane < nl) ? low.lists[(i64)li * CHAIN_LIST_LEN + lane] : -1;
        const i32 sv = e >= 0 ? ld_st(st, e) : 0;
        const u64 fm = __ballot(e >= 0 && sv < 0);
        const bool mine = (lane & ~(CH_SG - 1)) == l;
        if (fm) {
          const int r = __ffsll((unsigned long long)fm) - 1;
          const i32 r0 = __shfl(e, r), sr = __shfl(sv, r);
          if (mine) {
            spos = r;
            u = r0;
            s_u = sr;
            b = BLK_UNKNOWN;
            uknown = false;
            cpos = 0;
            pci = -1;
            ph = PK;
            pnv = 0;
          }
          need &= need - 1;
          continue;
        }
        if (mine) spos = CHAIN_LIST_LEN;
      }
      u64 kk;
      int nv;
      wave_best_free_topk<PK>(g, x, st, lane, l, kk, nv);
      const u64 key0 = __shfl(kk, l);
      const i64 xs0 = g.xadj[x];
      const i32 r0 = key0 ? g.adj[xs0 + (i64)(0xFFFFFFFFu - (u32)key0)] : -1;
      const i32 sr = r0 >= 0 ? ld_st(st, r0) : ST_UNDEC;
      if ((lane & ~(CH_SG - 1)) == l) {
        pci = kk ? g.adj[xs0 + (i64)(0xFFFFFFFFu - (u32)kk)] : -1;
        ph = 0;
        pnv = nv;
        u = r0 >= 0 ? r0 : PROP_NONE;
        s_u = sr;
        b = BLK_UNKNOWN;
        uknown = false;
        cpos = 0;
      }
      need &= need - 1;
    }
    // condition B: u above t, or u below t and decided unmatched
    bool okB = own && u >= 0 && (u > t || s_u == ST_EXHAUST);
    // condition C: no undecided neighbour of u below t (and none of them took u)
    if (okB && b >= 0 && s_b != ST_UNDEC) {
      if (s_b == u) {
        u = BLK_UNKNOWN;
        okB = false;
      } else {
        b = BLK_UNKNOWN;
      }
    }
    const bool needC = okB && b == BLK_UNKNOWN;
    const bool fetch = needC && !uknown;
    if (__ballot(fetch)) {  // the proposal's row, once per proposal
      if constexpr (STATS) c_fetch++;
      i64 us0 = 0, us1 = 0;
      if (fetch) {
        us0 = g.xadj[u];
        us1 = g.xadj[u + 1];
        ushort_ = us1 - us0 <= CH_ROW;
        uknown = true;
      }
#pragma unroll
      for (int j = 0; j < CH_PER; j++) {
        const i64 i = us0 + j * CH_SG + sl;
        if (fetch && ushort_) uv[j] = i < us1 ? g.adj[i] : -1;
      }
    }
    if (__ballot(needC && ushort_)) {
      if constexpr (STATS) c_cS++;
      i32 sv[CH_PER];
#pragma unroll
      for (int j = 0; j < CH_PER; j++)
        sv[j] = (needC && ushort_ && uv[j] >= 0 && uv[j] < t && uv[j] != u) ? ld_st(st, uv[j])
                                                                          : -4;
      i32 bestC = BLK_NONE;
      bool tk = false;
#pragma unroll
      for (int j = 0; j < CH_PER; j++) {
        if (sv[j] == ST_UNDEC) bestC = uv[j] > bestC ? uv[j] : bestC;
        tk |= sv[j] == u;
      }
      bestC = sg_max_i32<CH_SG>(bestC);
      tk = sg_any<CH_SG>(tk);
      if (needC && ushort_) {
        if (tk) {
          u = BLK_UNKNOWN;
          okB = false;
        } else {
          b = bestC;
        }
      }
    }
    need = __ballot(needC && !ushort_ && sl == 0);
    while (need) {
      const int l = __ffsll((unsigned long long)need) - 1;
      const i32 x = __shfl(u, l);
      const i32 lim = __shfl(t, l);
      i32 from = __shfl(cpos, l);
      // entries of u's row below t lie on u's lower-neighbour list when t <= u
      const bool uselow = low.off && lim <= x;
      const i32* la = uselow ? low.adj : g.adj;
      const i64 l0 = uselow ? low.off[x] : g.xadj[x], l1 = uselow ? low.off[x + 1] : g.xadj[x + 1];
      if constexpr (STATS) c_cL++;
      if constexpr (STATS) c_arcsC += l1 - l0 - from;
      const i32 r = wave_scan_pending(la, l0, l1, x, lim, x, st, lane, from, low.sorted);
      if ((lane & ~(CH_SG - 1)) == l) {
        if (r == BLK_TAKEN) {
          u = BLK_UNKNOWN;
          okB = false;
        } else {
          b = r;
        }
        cpos = from;
      }
      need &= need - 1;
    }
    const bool ready = own && a == BLK_NONE && (u == PROP_NONE || (okB && b == BLK_NONE));
    if (ready && sl == 0) {
      if (u >= 0) {
        st_mate(st, t, u);
        st_mate(st, u, t);
      } else {
        st_mate(st, t, ST_EXHAUST);
      }
    }
    if (ready) own = false;
    if constexpr (STATS) c_commit += __popcll(__ballot(ready && sl == 0));
    if (!__ballot(ready)) {
      if constexpr (STATS) c_sleep++;
      __builtin_amdgcn_s_sleep(2);
    }
  }
  if (own && sl == 0) {  // out of time: keep the caches for the next launch
    blkA[t] = a;
    prop[t] = u;
    blkC[t] = b;
  }
  if constexpr (STATS) {
    if (stats && lane == 0) {
      const u64 c[12] = {c_iter, c_sleep, c_apS, c_aL, c_pL, c_fetch, c_cS, c_cL, c_commit,
                         c_arcsL, c_arcsP, c_arcsC};
      for (int q = 0; q < 12; q++)
        atomicAdd((unsigned long long*)&stats[q], (unsigned long long)c[q]);
    }
  }
}

// STATS: the event counts of the chain (kept in registers only when asked)
template <int CH_SG, int CH_PER, bool STATS>
__global__ __launch_bounds__(256) void k_match_chain(DGraph g, ChainLow low, i32* st, i32* blkA,
                                                     i32* blkC, i32* prop, const i64* lo_io,
                                                     u32* ticket, i64 budget, u64* stats) {
  chain_wave<CH_SG, CH_PER, STATS>(g, low, st, blkA, blkC, prop, lo_io[0], ticket, budget,
                                   stats);
}

// Lower-neighbour lists of the rows longer than `minlen` (ChainLow).
// k_long_rows lists those rows (in no particular order); then, one wave per
// listed row with 16 loads per lane in flight, cnt[v] = number of entries
// u < v of row v, and, after an exclusive scan of cnt into off, the entries
// themselves in row order.
__global__ __launch_bounds__(256) void k_long_rows(DGraph g, i64 minlen, i32* list, i32* nlist) {
  // whole blocks iterate together, so every lane takes part in the ballot
  for (i64 b = (i64)blockIdx.x * blockDim.x; b < g.n; b += gstride()) {
    const i64 v = b + threadIdx.x;
    const bool lng = v < g.n && g.xadj[v + 1] - g.xadj[v] > minlen;
    wave_append(lng, (i32)v, list, nlist);
  }
}
template <bool WRITE>
__global__ __launch_bounds__(256) void k_chain_low(DGraph g, const i32* list, const i32* nlist,
                                                   i64* cnt, const i64* off, i32* out) {
  constexpr int U = 16;
  const int lane = lane_id();
  const i64 nl = *nlist;
  for (i64 e = gwave(); e < nl; e += nwaves()) {
    const i32 v = list[e];
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    i64 pos = WRITE ? off[v] : 0;
    for (i64 b = s0; b < s1; b += WAVE * U) {
      i32 u[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const i64 i = b + j * WAVE + lane;
        u[j] = i < s1 ? g.adj[i] : INT_MAX;
      }
#pragma unroll
      for (int j = 0; j < U; j++) {
        const bool f = u[j] < v;
        const u64 bm = __ballot(f);
        if (WRITE && f) out[pos + __popcll(bm & lanes_below(lane))] = u[j];
        pos += __popcll(bm);
      }
    }
    if (!WRITE && lane == 0) cnt[v] = pos;
  }
}

// st -> mate (and dec for the round kernels)
__global__ void k_chain_export(const i32* st, i64 n, i32* mate, u8* dec) {
  for (i64 i = gtid(); i < n; i += gstride()) {
    const i32 v = st[i];
    mate[i] = v >= 0 ? v : -1;
    dec[i] = v != ST_UNDEC;
  }
}

// lo_io[0] = first vertex at or above lo_io[0] with st == -1 (one block)
// *out = max over v of the row length xadj[v+1] - xadj[v] (zeroed by the caller)
__global__ __launch_bounds__(256) void k_max_row(DGraph g, u64* out) {
  u64 mx = 0;
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const u64 d = (u64)(g.xadj[v + 1] - g.xadj[v]);
    mx = d > mx ? d : mx;
  }
  mx = wave_max_u64(mx);
  if (lane_id() == 0 && mx) atomicMax((unsigned long long*)out, (unsigned long long)mx);
}

// First undecided vertex at or above lo[0]: blocks scan disjoint strips and
// fold their first hit into lo[3] (which holds n on entry) with atomicMin;
// k_first_undecided_take moves the result to lo[0] and resets lo[3] to n.
__global__ __launch_bounds__(256) void k_first_undecided_st(const i32* st, i64 n, i64* lo) {
  __shared__ i64 s_found;
  if (threadIdx.x == 0) s_found = LLONG_MAX;
  __syncthreads();
  const i64 start = lo[0];
  for (i64 i = start + gtid(); i < n; i += gstride()) {
    if (st[i] == ST_UNDEC) {
      atomicMin((unsigned long long*)&s_found, (unsigned long long)i);
      break;  // later hits of this thread are larger
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && s_found != LLONG_MAX)
    atomicMin((unsigned long long*)&lo[3], (unsigned long long)s_found);
}
__global__ void k_first_undecided_take(i64* lo, i64 n) {
  lo[0] = lo[3];
  lo[3] = n;
}

// Symmetry test of a level: every arc s->x (s != x, x < n) must have x->s.
// Also records each arc's weight and flags a repeated (v, u) as a duplicate,
// so the check also tells whether every row is simple and every arc's reverse
// has the same weight (flags[1], flags[2]).
__global__ __launch_bounds__(256) void k_sym_insert(DGraph g, u64* table, i32* wt, i64 cap,
                                                    i32* flags) {
  const u64 mask = (u64)cap - 1;
  const int lane = lane_id();
  bool dup = false;
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < 0 || u >= g.n || u == v) continue;
      const u64 key = ((u64)v << 32) | (u64)(u32)u;
      u64 h = hash64(key) & mask;
      while (true) {
        const u64 old = atomicCAS((unsigned long long*)&table[h], ~0ull, (unsigned long long)key);
        if (old == ~0ull) {
          wt[h] = g.w[i];
          break;
        }
        if (old == key) {
          dup = true;
          break;
        }
        h = (h + 1) & mask;
      }
    }
  }
  if (__ballot(dup) && lane == 0) atomicOr(&flags[1], 1);
}
__global__ __launch_bounds__(256) void k_sym_check(DGraph g, const u64* table, const i32* wt,
                                                   i64 cap, i32* flags) {
  const u64 mask = (u64)cap - 1;
  const int lane = lane_id();
  bool miss = false, wmis = false;
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < 0 || u >= g.n || u == v) continue;
      const u64 key = ((u64)(u32)u << 32) | (u64)v;
      u64 h = hash64(key) & mask;
      while (true) {
        const u64 cur = table[h];
        if (cur == key) {
          wmis |= wt[h] != g.w[i];
          break;
        }
        if (cur == ~0ull) {
          miss = true;
          break;
        }
        h = (h + 1) & mask;
      }
    }
  }
  if (__ballot(miss) && lane == 0) atomicOr(&flags[0], 1);
  if (__ballot(wmis) && lane == 0) atomicOr(&flags[2], 1);
}

// The shape of a CSR handed over in device memory: xadj[0] = 0, offsets
// non-decreasing, xadj[n] = m, every neighbour id in [0, n); *flag = 1 on any
// violation.  Reads the n + 1 offsets and the m ids only (within the arrays'
// declared sizes), so a malformed graph is rejected before any kernel follows
// its offsets.
__global__ void k_csr_check(const i64* xadj, const i32* adj, i64 n, i64 m, i32* flag) {
  bool bad = false;
  for (i64 v = gtid(); v <= n; v += gstride()) {
    const i64 x = xadj[v];
    if (v == 0) bad |= x != 0;
    if (v == n) bad |= x != m;
    else bad |= xadj[v + 1] < x;
  }
  for (i64 i = gtid(); i < m; i += gstride()) {
    const i32 u = adj[i];
    bad |= u < 0 || (i64)u >= n;
  }
  if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1);
}

// The same check for the device-resident graphs of a batch: one block per
// listed graph, bad[j] = 1 when graph j is malformed.
__global__ __launch_bounds__(256) void k_batch_csr_check(const BatchGraph* dg, const i32* list,
                                                         i32 cnt, i32* bad) {
  for (i32 e = blockIdx.x; e < cnt; e += gridDim.x) {
    const i32 j = list[e];
    const BatchGraph b = dg[j];
    bool f = false;
    for (i64 v = threadIdx.x; v <= b.n; v += blockDim.x) {
      const i64 x = b.xadj[v];
      if (v == 0) f |= x != 0;
      if (v == b.n) f |= x != b.m;
      else f |= b.xadj[v + 1] < x;
    }
    for (i64 i = threadIdx.x; i < b.m; i += blockDim.x) {
      const i32 u = b.adj[i];
      f |= u < 0 || (i64)u >= b.n;
    }
    if (__ballot(f) && lane_id() == 0) atomicOr(&bad[j], 1);
  }
}

// Bit length of each row's degree (the rank of a row in the search below).
__global__ void k_deg_class(DGraph g, u8* cls) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 d = g.xadj[v + 1] - g.xadj[v];
    cls[v] = d > 0 ? (u8)(64 - __clzll((unsigned long long)d)) : 0;
  }
}

// crow[c] = the row holding entry c * SYM_CHUNK (chunks of the arc range)
constexpr int SYM_CHUNK = 1024;
__global__ void k_chunk_rows(DGraph g, i64 nch, i32* crow) {
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    if (s1 <= s0) continue;
    for (i64 c = (s0 + SYM_CHUNK - 1) / SYM_CHUNK; c <= (s1 - 1) / SYM_CHUNK && c < nch; c++)
      crow[c] = (i32)v;
  }
}

// flags[3] |= some row is not strictly ascending; flags[5] |= some arc between
// distinct vertices in range weighs other than 1.  One wave per chunk of
// SYM_CHUNK arcs (crow[c] = the row of chunk c's first arc), 64 arcs per step:
// each lane finds its arc's row by binary search over the ends of the 64 rows
// from the current one, so short rows share a step instead of a wave each.
__global__ __launch_bounds__(256) void k_rows_strict(DGraph g, const i32* crow, i64 nch, i64 m,
                                                     i32* flags) {
  const int lane = lane_id();
  bool bad = false, nonunit = false;
  for (i64 c = gwave(); c < nch; c += nwaves()) {
    i64 b = c * SYM_CHUNK;
    const i64 bend = b + SYM_CHUNK < m ? b + SYM_CHUNK : m;
    i64 r = crow[c];
    i64 rs = g.xadj[r];  // start of row r
    while (b < bend) {
      const i64 rl = r + lane;
      const i64 e = g.xadj[rl < g.n ? rl + 1 : g.n];  // end of row rl
      // rows of the window that end at or before b hold none of the entries
      const int done = __popcll(__ballot(e <= b));
      if (done) {
        rs = __shfl(e, done - 1);
        r += done;
        continue;
      }
      const i64 e63 = __shfl(e, WAVE - 1);
      const i64 lim = b + WAVE < bend ? b + WAVE : bend;
      const i64 wend = lim < e63 ? lim : e63;
      const i64 i = b + lane;
      int rr = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(e, rr + step - 1) <= i) rr += step;
      // start of the row of i (the shuffle runs on every lane: a lane left
      // out of it would read 0 where it is the source)
      const i64 ep = __shfl(e, rr > 0 ? rr - 1 : 0);
      const i64 start = rr ? ep : rs;
      b = wend;
      if (i >= wend) continue;
      const i64 v = r + rr;
      const i32 u = g.adj[i];
      if (i > start) bad |= u <= g.adj[i - 1];
      if (u >= 0 && u < g.n && u != v) nonunit |= g.w[i] != 1;
    }
  }
  if (__ballot(bad) && lane == 0) atomicOr(&flags[3], 1);
  if (__ballot(nonunit) && lane == 0) atomicOr(&flags[5], 1);
}

// Strictly ascending rows.  Of the two arcs of a pair {u, v}, the one whose
// target row ranks lower by (degree bit length, id) binary-searches that row
// for its source and compares weights (counted in cnt[1] when found, flags[0]
// when missing, flags[2] on a weight mismatch); the other is counted in
// cnt[0].  Rows hold no repeats, so distinct searching arcs find distinct
// counted arcs: symmetric iff nothing is missing and both counts agree.  The
// searched row is never much longer than the searching one, so on skewed
// degrees the searches stay short.  One wave per chunk of SYM_CHUNK entries
// of the arc range (long rows spread over many waves): 64 entries per step,
// each lane finding its row by binary search over the ends of the 64 rows
// from the current one (held one per lane).  flags[4] is set when a found
// pair carries a weight other than 1 (on a symmetric level the found pairs
// hold every arc between distinct vertices in range).
__global__ __launch_bounds__(256) void k_sym_search_sorted(DGraph g, const u8* cls,
                                                           const i32* crow, i64 nch, i64 m,
                                                           i32* flags, u64* cnt) {
  const int lane = lane_id();
  bool miss = false, wmis = false, nonunit = false;
  u64 cu = 0, cd = 0;
  for (i64 c = gwave(); c < nch; c += nwaves()) {
    i64 b = c * SYM_CHUNK;
    const i64 bend = b + SYM_CHUNK < m ? b + SYM_CHUNK : m;
    i64 r = crow[c];
    while (b < bend) {
      const i64 rl = r + lane;
      const i64 e = g.xadj[rl < g.n ? rl + 1 : g.n];  // end of row rl
      const i32 cl = rl < g.n ? cls[rl] : 0;
      const i64 i = b + lane;
      const i32 u = i < bend ? g.adj[i] : -1;
      // rows of the window that end at or before b hold none of the entries
      const int done = __popcll(__ballot(e <= b));
      if (done) {
        r += done;
        continue;
      }
      const i64 e63 = __shfl(e, WAVE - 1);
      const i64 lim = b + WAVE < bend ? b + WAVE : bend;
      const i64 wend = lim < e63 ? lim : e63;
      // lanes whose row ends at or before i: the row of i is r + that count
      int rr = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1)
        if (__shfl(e, rr + step - 1) <= i) rr += step;
      const i32 cv = __shfl(cl, rr);
      const i64 v = r + rr;
      b = wend;
      if (i >= wend || u >= g.n || u < 0 || u == v) continue;
      const i32 cuu = cls[u];
      if (!(cuu < cv || (cuu == cv && u < v))) {
        cu++;
        continue;
      }
      i64 lo = g.xadj[u], hi = g.xadj[u + 1];  // first entry >= v in [lo, hi)
      const i64 end = hi;
      while (lo < hi) {
        const i64 mid = lo + ((hi - lo) >> 1);
        if (g.adj[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      if (lo < end && g.adj[lo] == v) {
        const i32 wl = g.w[lo], wi = g.w[i];
        wmis |= wl != wi;
        nonunit |= wl != 1 || wi != 1;
        cd++;
      } else {
        miss = true;
      }
    }
  }
  cu = wave_sum(cu);
  cd = wave_sum(cd);
  if (lane == 0 && cu) atomicAdd((unsigned long long*)&cnt[0], (unsigned long long)cu);
  if (lane == 0 && cd) atomicAdd((unsigned long long*)&cnt[1], (unsigned long long)cd);
  if (__ballot(miss) && lane == 0) atomicOr(&flags[0], 1);
  if (__ballot(wmis) && lane == 0) atomicOr(&flags[2], 1);
  if (__ballot(nonunit) && lane == 0) atomicOr(&flags[4], 1);
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

// Row classes by enumerated size d (fine row + mate row): d <= REP_SMALL rows
// are deduplicated in registers, d <= REP_MED rows in an LDS table, longer rows
// in global tables (dlong = d and csz = table capacity only for those).
constexpr i64 REP_SMALL = 64;
constexpr i64 REP_MED = 2048;
// medium rows of at most REP_MEDW entries take one wave each (medS)
constexpr i64 REP_MEDW = 256;

__global__ void k_rep_classify(const i64* xadj, const i32* tb, i64 n_c, i64* dsz, i64* dlong,
                               i64* csz, i32* med, i32* nmed, i32* lng, i32* nlng, i64* nchk,
                               i32* meds, i32* nmeds) {
  for (i64 c = gtid(); c < n_c; c += gstride()) {
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    i64 d = xadj[r + 1] - xadj[r];
    if (mt >= 0) d += xadj[mt + 1] - xadj[mt];
    dsz[c] = d;
    const bool is_long = d > REP_MED;
    dlong[c] = d > REP_SMALL ? d : 0;
    csz[c] = is_long ? next_pow2_i64(2 * d) : 0;
    if (is_long) {
      const i32 e = atomicAdd(nlng, 1);
      lng[e] = (i32)c;
      nchk[e] = (d + REP_CHUNK_ENTRIES - 1) / REP_CHUNK_ENTRIES;
    }
    wave_append(!is_long && d > REP_MEDW, (i32)c, med, nmed);
    wave_append(d > REP_SMALL && d <= REP_MEDW, (i32)c, meds, nmeds);
  }
}

// Long rows (d > REP_MED): chunks of REP_CHUNK entries, one block per chunk,
// sharing the row's global table.  A table slot is one 64-bit word
// (x << 32 | first position), so an insert is one compare-and-swap, or one
// 64-bit atomicMin when x is already there; weights go to W by atomicAdd.
constexpr i64 REP_CHUNK = REP_CHUNK_ENTRIES;
constexpr u64 REP_EMPTY = ~0ull;

// chunk descriptors of the long rows: row index in the list, first entry
__global__ void k_rep_chunks(const i32* lng, i64 nl, const i64* dsz, const i64* cbase,
                             i32* ch_row, i64* ch_beg) {
  for (i64 e = gtid(); e < nl; e += gstride()) {
    const i64 d = dsz[lng[e]];
    i64 o = cbase[e];
    for (i64 b = 0; b < d; b += REP_CHUNK, o++) {
      ch_row[o] = (i32)e;
      ch_beg[o] = b;
    }
  }
}

__device__ __forceinline__ void rep_fetch(const DGraph& g, const i32* tb, const i32* new_ids,
                                          i64 c, i64 p, i32* x, i32* w) {
  const i32 r = tb[2 * c], mt = tb[2 * c + 1];
  const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
  const i64 i = p < la ? a0 + p : g.xadj[mt] + (p - la);
  *x = new_ids[g.adj[i]];
  *w = g.w[i];
}

__global__ __launch_bounds__(256) void k_rep_insert(DGraph g, const i32* tb, const i32* new_ids,
                                                    const i32* lng, const i32* ch_row,
                                                    const i64* ch_beg, i64 nch, const i64* dsz,
                                                    const i64* toff, const i64* csz, u64* T,
                                                    i32* W) {
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i64 c = lng[ch_row[k2]];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 e = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const u64 mask = (u64)(csz[c] - 1);
    u64* Tc = T + toff[c];
    i32* Wc = W + toff[c];
    for (i64 p = b + threadIdx.x; p < e; p += blockDim.x) {
      i32 x, w;
      rep_fetch(g, tb, new_ids, c, p, &x, &w);
      if (x == (i32)c || x == -1) continue;
      const u64 word = ((u64)(u32)x << 32) | (u64)(u32)p;
      u64 h = (u64)hash32((u32)x) & mask;
      while (true) {
        const u64 old = atomicCAS((unsigned long long*)&Tc[h], (unsigned long long)REP_EMPTY,
                                  (unsigned long long)word);
        if (old == REP_EMPTY) break;
        if ((old >> 32) == (u64)(u32)x) {
          if ((u32)old > (u32)p) atomicMin((unsigned long long*)&Tc[h], (unsigned long long)word);
          break;
        }
        h = (h + 1) & mask;
      }
      atomicAdd(&Wc[h], w);
    }
  }
}

// E[p] = table slot of entry p when it is its x's first occurrence, else -1;
// chunk_cnt = number of first occurrences in the chunk.
__global__ __launch_bounds__(256) void k_rep_mark(DGraph g, const i32* tb, const i32* new_ids,
                                                  const i32* lng, const i32* ch_row,
                                                  const i64* ch_beg, i64 nch, const i64* dsz,
                                                  const i64* toff, const i64* csz,
                                                  const i64* eoff, const u64* T, i32* E,
                                                  i64* chunk_cnt) {
  __shared__ i64 s_c[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i64 c = lng[ch_row[k2]];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 e = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const u64 mask = (u64)(csz[c] - 1);
    const u64* Tc = T + toff[c];
    i32* Ec = E + eoff[c];
    i64 reps = 0;
    for (i64 p = b + threadIdx.x; p < e; p += blockDim.x) {
      i32 x, w;
      rep_fetch(g, tb, new_ids, c, p, &x, &w);
      i32 sl = -1;
      if (!(x == (i32)c || x == -1)) {
        u64 h = (u64)hash32((u32)x) & mask;
        u64 t;
        while (((t = Tc[h]) >> 32) != (u64)(u32)x) h = (h + 1) & mask;
        if ((u32)t == (u32)p) sl = (i32)h;
      }
      Ec[p] = sl;
      reps += sl >= 0;
    }
    reps = wave_sum(reps);
    if (lane == 0) s_c[wv] = reps;
    __syncthreads();
    if (threadIdx.x == 0) {
      i64 t2 = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) t2 += s_c[w2];
      chunk_cnt[k2] = t2;
    }
    __syncthreads();
  }
}

// cnt[c] of every long row = sum of its chunks' counts (chunks are consecutive)
__global__ void k_rep_rowcnt(const i32* lng, i64 nl, const i64* cbase, const i64* choff,
                             i64* cnt) {
  for (i64 e = gtid(); e < nl; e += gstride()) cnt[lng[e]] = choff[cbase[e + 1]] - choff[cbase[e]];
}

// writes the first occurrences of a chunk at ox[c] + (its earlier chunks' counts)
__global__ __launch_bounds__(256) void k_rep_emit(const i32* lng, const i32* ch_row,
                                                  const i64* ch_beg, i64 nch, const i64* dsz,
                                                  const i64* cbase, const i64* choff,
                                                  const i64* toff, const i64* eoff, const u64* T,
                                                  const i32* W, const i32* E, const i64* ox,
                                                  i32* oa, i32* ow) {
  __shared__ i32 s_w[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i32 e0 = ch_row[k2];
    const i64 c = lng[e0];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 e = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const u64* Tc = T + toff[c];
    const i32* Wc = W + toff[c];
    const i32* Ec = E + eoff[c];
    i64 run = ox[c] + (choff[k2] - choff[cbase[e0]]);
    for (i64 base = b; base < e; base += blockDim.x) {
      const i64 p = base + threadIdx.x;
      const i32 sl = p < e ? Ec[p] : -1;
      const u64 bm = __ballot(sl >= 0);
      if (lane == 0) s_w[wv] = __popcll(bm);
      __syncthreads();
      i64 before = 0, tot = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) {
        if (w2 < wv) before += s_w[w2];
        tot += s_w[w2];
      }
      if (sl >= 0) {
        const i64 o = run + before + __popcll(bm & lanes_below(lane));
        oa[o] = (i32)(Tc[sl] >> 32);
        ow[o] = Wc[sl];
      }
      run += tot;
      __syncthreads();
    }
  }
}

// Long rows, bucketed form (the default; the global tables above are the
// fallback).  Every entry of a long row gets a word R[p] = x << 32 | w in
// position order (x = ~0 for a self loop).  Entries are then scattered into
// 2^lg buckets of the row by the top lg bits of hash32(x) (about RL_BK_TARGET
// entries each, or the target the host passes), and one block per bucket finds, in an LDS table, each
// coarse neighbour's first position and weight sum.  Only neighbours that
// occur more than once change R: the first occurrence gets the summed
// weight, the others are dropped.  The row is then compacted in position
// order from R.  A bucket whose distinct neighbours would fill more than
// RL_TFILL of its table stops the bucketed form for the whole level (the
// host reruns the level's long rows through the global tables).
constexpr int RL_BK_LGMAX = 13;  // at most 8192 buckets per row (LDS histogram)
constexpr i64 RL_BK_TARGET = 1024;
constexpr int RL_TCAP = 4096;
constexpr int RL_TFILL = REP_BUCKET_FILL;
static_assert(RL_TFILL + 256 < RL_TCAP, "a free slot must remain after the fill limit");
constexpr u32 RL_NONE = 0xFFFFFFFFu;
constexpr int RL_PER = (int)(REP_CHUNK / 256);  // entries per thread of a chunk

__device__ __forceinline__ u32 rl_bucket(u32 x, int lg) {
  return lg ? hash32(x) >> (32 - lg) : 0u;
}

// lgb[e] = log2 of the bucket count of long row e; nbk[e] = 2^lgb[e]
__global__ void k_rl_buckets(const i32* lng, i64 nl, const i64* dsz, i64 target, i32* lgb,
                             i64* nbk) {
  for (i64 e = gtid(); e < nl; e += gstride()) {
    const i64 want = (dsz[lng[e]] + target - 1) / target;
    int lg = 0;
    while (lg < RL_BK_LGMAX && ((i64)1 << lg) < want) lg++;
    lgb[e] = lg;
    nbk[e] = (i64)1 << lg;
  }
}
// brow[b] = the long row that bucket b belongs to
__global__ void k_rl_brow(i64 nl, const i64* bbase, i32* brow) {
  for (i64 e = gtid(); e < nl; e += gstride())
    for (i64 b = bbase[e]; b < bbase[e + 1]; b++) brow[b] = (i32)e;
}

// R words of a chunk, and its per-bucket entry counts added to bcnt
__global__ __launch_bounds__(256) void k_rl_fetch(DGraph g, const i32* tb, const i32* new_ids,
                                                  const i32* lng, const i32* ch_row,
                                                  const i64* ch_beg, i64 nch, const i64* dsz,
                                                  const i64* eoff, const i32* lgb,
                                                  const i64* bbase, u64* R, u64* bcnt) {
  __shared__ i32 H[1 << RL_BK_LGMAX];
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i32 e0 = ch_row[k2];
    const i64 c = lng[e0];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 end = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const int lg = lgb[e0];
    const i32 nb = 1 << lg;
    for (i32 h = threadIdx.x; h < nb; h += blockDim.x) H[h] = 0;
    __syncthreads();
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0;
    u64* Rc = R + eoff[c];
    i32 v[RL_PER], w[RL_PER];
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const i64 p = b + j * 256 + threadIdx.x;
      const i64 i = p < la ? a0 + p : b0 + (p - la);
      v[j] = p < end ? g.adj[i] : -1;
      w[j] = p < end ? g.w[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < RL_PER; j++) v[j] = v[j] >= 0 ? new_ids[v[j]] : -1;
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const i64 p = b + j * 256 + threadIdx.x;
      if (p < end) {
        const bool keep = !(v[j] == (i32)c || v[j] == -1);
        Rc[p] = keep ? (((u64)(u32)v[j] << 32) | (u64)(u32)w[j]) : ((u64)RL_NONE << 32);
        if (keep) atomicAdd(&H[rl_bucket((u32)v[j], lg)], 1);
      }
    }
    __syncthreads();
    const i64 bb = bbase[e0];
    for (i32 h = threadIdx.x; h < nb; h += blockDim.x)
      if (H[h]) atomicAdd((unsigned long long*)&bcnt[bb + h], (unsigned long long)H[h]);
    __syncthreads();
  }
}

// the entries of a chunk into their buckets: K = x << 32 | position, W = weight
// (order within a bucket is arbitrary; the bucket's result does not depend on it)
__global__ __launch_bounds__(256) void k_rl_scatter(const i32* lng, const i32* ch_row,
                                                    const i64* ch_beg, i64 nch, const i64* dsz,
                                                    const i64* eoff, const i32* lgb,
                                                    const i64* bbase, const i64* boff,
                                                    u32* bcur, const u64* R, u64* K, i32* W) {
  __shared__ i32 H[1 << RL_BK_LGMAX];
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i32 e0 = ch_row[k2];
    const i64 c = lng[e0];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 end = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const int lg = lgb[e0];
    const i32 nb = 1 << lg;
    const i64 bb = bbase[e0];
    for (i32 h = threadIdx.x; h < nb; h += blockDim.x) H[h] = 0;
    __syncthreads();
    const u64* Rc = R + eoff[c];
    u64 rv[RL_PER];
    i32 rk[RL_PER];
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const i64 p = b + j * 256 + threadIdx.x;
      rv[j] = p < end ? Rc[p] : ((u64)RL_NONE << 32);
    }
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const u32 x = (u32)(rv[j] >> 32);
      rk[j] = x != RL_NONE ? atomicAdd(&H[rl_bucket(x, lg)], 1) : -1;
    }
    __syncthreads();
    // counts become this chunk's first slot within each bucket
    for (i32 h = threadIdx.x; h < nb; h += blockDim.x)
      if (H[h]) H[h] = (i32)atomicAdd(&bcur[bb + h], (u32)H[h]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      if (rk[j] < 0) continue;
      const u32 x = (u32)(rv[j] >> 32);
      const u32 bk = rl_bucket(x, lg);
      const i64 o = boff[bb + bk] + H[bk] + rk[j];
      const i64 p = b + j * 256 + threadIdx.x;
      K[o] = ((u64)x << 32) | (u64)(u32)p;
      W[o] = (i32)(u32)rv[j];
    }
    __syncthreads();
  }
}

// One block per bucket: first position, weight sum and multiplicity of each
// neighbour in an LDS table; R of repeated neighbours rewritten.  Records are
// taken 2048 at a time (8 per thread); a bucket of at most 2048 records keeps
// them in registers between the two passes.
constexpr int RL_SLICE = 2048;
__device__ __forceinline__ void rl_insert(const u64* kk, const i32* ww, i32* TK, i32* TP,
                                          i32* TW, u8* TD, i32* s_n, i32* s_over, i32 fill) {
#pragma unroll
  for (int j = 0; j < RL_SLICE / 256; j++) {
    // once the table holds `fill` keys nothing more is inserted: at most one
    // insert per thread is then still in flight, so a free slot always remains
    // (fill + 256 < RL_TCAP) and every probe ends
    if (kk[j] == ~0ull || *(volatile i32*)s_over) continue;
    const i32 x = (i32)(kk[j] >> 32), p = (i32)(u32)kk[j];
    u32 h = hash32((u32)x) & (RL_TCAP - 1);
    while (true) {
      const i32 old = atomicCAS(&TK[h], -1, x);
      if (old == -1) {
        if (atomicAdd(s_n, 1) >= fill) *s_over = 1;
        break;
      }
      if (old == x) {
        TD[h] = 1;
        break;
      }
      h = (h + 1) & (RL_TCAP - 1);
    }
    atomicMin(&TP[h], p);
    atomicAdd(&TW[h], ww[j]);
  }
}
__device__ __forceinline__ void rl_apply(const u64* kk, const i32* TK, const i32* TP,
                                         const i32* TW, const u8* TD, u64* Rc) {
#pragma unroll
  for (int j = 0; j < RL_SLICE / 256; j++) {
    if (kk[j] == ~0ull) continue;
    const i32 x = (i32)(kk[j] >> 32), p = (i32)(u32)kk[j];
    u32 h = hash32((u32)x) & (RL_TCAP - 1);
    while (TK[h] != x) h = (h + 1) & (RL_TCAP - 1);
    if (!TD[h]) continue;
    Rc[p] = TP[h] == p ? (((u64)(u32)x << 32) | (u64)(u32)TW[h]) : ((u64)RL_NONE << 32);
  }
}
__device__ __forceinline__ void rl_load(const u64* K, const i32* W, i64 base, i64 r1, u64* kk,
                                        i32* ww) {
#pragma unroll
  for (int j = 0; j < RL_SLICE / 256; j++) {
    const i64 i = base + j * 256 + threadIdx.x;
    kk[j] = i < r1 ? K[i] : ~0ull;
    ww[j] = i < r1 ? W[i] : 0;
  }
}
__global__ __launch_bounds__(256) void k_rl_dedup(const i64* boff, i64 nbt, const i32* brow,
                                                  const i32* lng, const i64* eoff, const u64* K,
                                                  const i32* W, u64* R, i32 fill, i32* over) {
  __shared__ i32 TK[RL_TCAP], TP[RL_TCAP], TW[RL_TCAP];
  __shared__ u8 TD[RL_TCAP];
  __shared__ i32 s_n, s_over;
  for (i64 bk = blockIdx.x; bk < nbt; bk += gridDim.x) {
    const i64 r0 = boff[bk], r1 = boff[bk + 1];
    if (r1 == r0) continue;  // uniform over the block
    u64* Rc = R + eoff[lng[brow[bk]]];
    for (int h = threadIdx.x; h < RL_TCAP; h += blockDim.x) {
      TK[h] = -1;
      TP[h] = INT_MAX;
      TW[h] = 0;
      TD[h] = 0;
    }
    if (threadIdx.x == 0) {
      s_n = 0;
      s_over = 0;
    }
    __syncthreads();
    u64 kk[RL_SLICE / 256];
    i32 ww[RL_SLICE / 256];
    const bool one = r1 - r0 <= RL_SLICE;
    bool ov = false;
    for (i64 base = r0; base < r1; base += RL_SLICE) {
      rl_load(K, W, base, r1, kk, ww);
      rl_insert(kk, ww, TK, TP, TW, TD, &s_n, &s_over, fill);
      __syncthreads();
      ov = s_over != 0;
      __syncthreads();  // every thread has read s_over before the next slice may set it
      if (ov) break;
    }
    if (ov) {
      if (threadIdx.x == 0) atomicOr(over, 1);
      __syncthreads();
      continue;
    }
    if (one) {
      rl_apply(kk, TK, TP, TW, TD, Rc);
    } else {
      for (i64 base = r0; base < r1; base += RL_SLICE) {
        rl_load(K, W, base, r1, kk, ww);
        rl_apply(kk, TK, TP, TW, TD, Rc);
      }
    }
    __syncthreads();
  }
}

// entries left in R per chunk (chunk_cnt), then the chunk's entries written in
// position order at ox[c] + (the earlier chunks' counts)
__global__ __launch_bounds__(256) void k_rl_count(const i32* lng, const i32* ch_row,
                                                  const i64* ch_beg, i64 nch, const i64* dsz,
                                                  const i64* eoff, const u64* R,
                                                  i64* chunk_cnt) {
  __shared__ i32 s_c[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i64 c = lng[ch_row[k2]];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 end = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const u64* Rc = R + eoff[c];
    i32 cnt = 0;
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const i64 p = b + j * 256 + threadIdx.x;
      cnt += (p < end && (u32)(Rc[p] >> 32) != RL_NONE) ? 1 : 0;
    }
    cnt = wave_sum(cnt);
    if (lane == 0) s_c[wv] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      i64 t = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) t += s_c[w2];
      chunk_cnt[k2] = t;
    }
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_rl_emit(const i32* lng, const i32* ch_row,
                                                 const i64* ch_beg, i64 nch, const i64* dsz,
                                                 const i64* cbase, const i64* choff,
                                                 const i64* eoff, const u64* R, const i64* ox,
                                                 i32* oa, i32* ow) {
  __shared__ i32 s_w[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 k2 = blockIdx.x; k2 < nch; k2 += gridDim.x) {
    const i32 e0 = ch_row[k2];
    const i64 c = lng[e0];
    const i64 d = dsz[c];
    const i64 b = ch_beg[k2];
    const i64 end = b + REP_CHUNK < d ? b + REP_CHUNK : d;
    const u64* Rc = R + eoff[c];
    u64 rv[RL_PER];
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const i64 p = b + j * 256 + threadIdx.x;
      rv[j] = p < end ? Rc[p] : ((u64)RL_NONE << 32);
    }
    i64 run = ox[c] + (choff[k2] - choff[cbase[e0]]);
#pragma unroll
    for (int j = 0; j < RL_PER; j++) {
      const bool keep = (u32)(rv[j] >> 32) != RL_NONE;
      const u64 bm = __ballot(keep);
      if (lane == 0) s_w[wv] = __popcll(bm);
      __syncthreads();
      i64 before = 0, tot = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) {
        if (w2 < wv) before += s_w[w2];
        tot += s_w[w2];
      }
      if (keep) {
        const i64 o = run + before + __popcll(bm & lanes_below(lane));
        oa[o] = (i32)(rv[j] >> 32);
        ow[o] = (i32)(u32)rv[j];
      }
      run += tot;
      __syncthreads();
    }
  }
}

// Rows with d <= SG entries: SG lanes per row, one entry per lane.  An entry
// is the first occurrence of its coarse neighbour x when no earlier lane holds
// x; its weight is the sum over all lanes holding x.  PASS 0 writes the row's
// count, PASS 1 writes the first occurrences, in position order, at ox[c].
template <int SG, int PASS>
__global__ __launch_bounds__(256) void k_rep_small(DGraph g, const i32* tb, const i32* new_ids,
                                                   i64 n_c, const i64* dsz, i64 dmin, i64* cnt,
                                                   const i64* ox, i32* oa, i32* ow) {
  const int lane = lane_id();
  const int sl = lane & (SG - 1);
  const int sgb = lane & ~(SG - 1);
  const i64 nsg = (i64)gridDim.x * (blockDim.x / SG);
  for (i64 c = (i64)blockIdx.x * (blockDim.x / SG) + threadIdx.x / SG; c < n_c; c += nsg) {
    const i64 d = dsz[c];
    if (d < dmin || d > SG) continue;
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0;
    bool valid = sl < d;
    i32 x = -1, w = 0;
    if (valid) {
      const i64 i = sl < la ? a0 + sl : b0 + (sl - la);
      x = new_ids[g.adj[i]];
      w = g.w[i];
      valid = !(x == (i32)c || x == -1);
    }
    bool first = valid;
    u32 wsum = 0;
    const int dq = (int)d;
    for (int q = 0; q < (SG == WAVE ? dq : SG); q++) {
      const i32 xq = __shfl(x, q, SG);
      const i32 wq = __shfl(w, q, SG);
      const int vq = __shfl((int)valid, q, SG);
      if (vq && valid && xq == x) {
        if (q < sl) first = false;
        wsum += (u32)wq;
      }
    }
    const u64 fm = (__ballot(first) >> sgb) & (SG == WAVE ? ~0ull : ((1ull << SG) - 1));
    if (PASS == 0) {
      if (sl == 0) cnt[c] = __popcll(fm);
    } else if (first) {
      const i64 o = ox[c] + __popcll(fm & ((1ull << sl) - 1));
      oa[o] = x;
      ow[o] = (i32)wsum;
    }
  }
}

// Rows with REP_SMALL < d <= REP_MED: one block per row.  Pass 0 loads the
// row's entries (8 per thread, all loads issued before the gathers), records
// each coarse neighbour's first position and weight sum in an LDS table of
// capacity pow2 >= 2d, writes every entry's word R[eoff[c] + p] (x << 32 |
// summed weight for a first occurrence, x = ~0 otherwise) and the row's count;
// pass 1 writes the kept words in position order (block-wide ranks over
// 256-entry steps) without touching the fine rows again.
constexpr int REP_TCAP = 4096;
constexpr int REP_MED_PER = (int)(REP_MED / 256);
__global__ __launch_bounds__(256) void k_rep_med0(DGraph g, const i32* tb, const i32* new_ids,
                                                  const i32* lst, i64 nl, const i64* dsz,
                                                  const i64* eoff, u64* R, i64* cnt) {
  __shared__ i32 K[REP_TCAP], P[REP_TCAP], W[REP_TCAP];
  __shared__ i32 s_w[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 e = blockIdx.x; e < nl; e += gridDim.x) {
    const i64 c = lst[e];
    const i64 d = dsz[c];
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0;
    const i32 cap = (i32)next_pow2_i64(2 * d);
    const u32 mask = (u32)(cap - 1);
    for (int h = threadIdx.x; h < cap; h += blockDim.x) {
      K[h] = -1;
      P[h] = INT_MAX;
      W[h] = 0;
    }
    i32 x[REP_MED_PER], w[REP_MED_PER];
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      const i64 p = j * 256 + threadIdx.x;
      const i64 i = p < la ? a0 + p : b0 + (p - la);
      x[j] = p < d ? g.adj[i] : -1;
      w[j] = p < d ? g.w[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      x[j] = x[j] >= 0 ? new_ids[x[j]] : -1;
      if (x[j] == (i32)c) x[j] = -1;
    }
    __syncthreads();  // table cleared
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      if (x[j] < 0) continue;
      u32 h = hash32((u32)x[j]) & mask;
      while (true) {
        const i32 old = atomicCAS(&K[h], -1, x[j]);
        if (old == -1 || old == x[j]) break;
        h = (h + 1) & mask;
      }
      atomicMin(&P[h], j * 256 + (i32)threadIdx.x);
      atomicAdd(&W[h], w[j]);
    }
    __syncthreads();
    u64* Rc = R + eoff[c];
    i32 kept = 0;
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      const i64 p = j * 256 + threadIdx.x;
      if (p >= d) continue;
      u64 word = (u64)RL_NONE << 32;
      if (x[j] >= 0) {
        u32 h = hash32((u32)x[j]) & mask;
        while (K[h] != x[j]) h = (h + 1) & mask;
        if (P[h] == (i32)p) {
          word = ((u64)(u32)x[j] << 32) | (u64)(u32)W[h];
          kept++;
        }
      }
      Rc[p] = word;
    }
    kept = wave_sum(kept);
    if (lane == 0) s_w[wv] = kept;
    __syncthreads();
    if (threadIdx.x == 0) {
      i64 t = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) t += s_w[w2];
      cnt[c] = t;
    }
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_rep_med1(const i32* lst, i64 nl, const i64* dsz,
                                                  const i64* eoff, const u64* R, const i64* ox,
                                                  i32* oa, i32* ow) {
  __shared__ i32 s_w[256 / WAVE];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 e = blockIdx.x; e < nl; e += gridDim.x) {
    const i64 c = lst[e];
    const i64 d = dsz[c];
    const u64* Rc = R + eoff[c];
    u64 rv[REP_MED_PER];
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      const i64 p = j * 256 + threadIdx.x;
      rv[j] = p < d ? Rc[p] : ((u64)RL_NONE << 32);
    }
    i64 run = ox[c];
#pragma unroll
    for (int j = 0; j < REP_MED_PER; j++) {
      if (j * 256 >= d) break;  // uniform over the block
      const bool keep = (u32)(rv[j] >> 32) != RL_NONE;
      const u64 bm = __ballot(keep);
      if (lane == 0) s_w[wv] = __popcll(bm);
      __syncthreads();
      i64 before = 0, tot = 0;
      for (int w2 = 0; w2 < 256 / WAVE; w2++) {
        if (w2 < wv) before += s_w[w2];
        tot += s_w[w2];
      }
      if (keep) {
        const i64 o = run + before + __popcll(bm & lanes_below(lane));
        oa[o] = (i32)(rv[j] >> 32);
        ow[o] = (i32)(u32)rv[j];
      }
      run += tot;
      __syncthreads();
    }
  }
}

// Medium rows of at most REP_MEDW entries: k_rep_med0/1 with one wave per
// row (4 entries per lane, a 512-slot LDS table per wave, no block barriers:
// a wave's LDS accesses are ordered by LDS-scope fences).
constexpr int REP_MEDW_PER = (int)(REP_MEDW / WAVE);
__device__ __forceinline__ void lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup", "local");
}
__global__ __launch_bounds__(256) void k_rep_medw0(DGraph g, const i32* tb, const i32* new_ids,
                                                   const i32* lst, i64 nl, const i64* dsz,
                                                   const i64* eoff, u64* R, i64* cnt) {
  constexpr int TC = (int)(2 * REP_MEDW);
  __shared__ i32 Ks[256 / WAVE][TC], Ps[256 / WAVE][TC], Ws[256 / WAVE][TC];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  i32* K = Ks[wv];
  i32* P = Ps[wv];
  i32* W = Ws[wv];
  for (i64 e = gwave(); e < nl; e += nwaves()) {
    const i64 c = lst[e];
    const i64 d = dsz[c];
    const i32 r = tb[2 * c], mt = tb[2 * c + 1];
    const i64 a0 = g.xadj[r], la = g.xadj[r + 1] - a0;
    const i64 b0 = mt >= 0 ? g.xadj[mt] : 0;
    const i32 cap = (i32)next_pow2_i64(2 * d);
    const u32 mask = (u32)(cap - 1);
    i32 x[REP_MEDW_PER], w[REP_MEDW_PER];
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      const i64 p = j * WAVE + lane;
      const i64 i = p < la ? a0 + p : b0 + (p - la);
      x[j] = p < d ? g.adj[i] : -1;
      w[j] = p < d ? g.w[i] : 0;
    }
    lds_order();  // the previous row's lookups are done
    for (int h = lane; h < cap; h += WAVE) {
      K[h] = -1;
      P[h] = INT_MAX;
      W[h] = 0;
    }
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      x[j] = x[j] >= 0 ? new_ids[x[j]] : -1;
      if (x[j] == (i32)c) x[j] = -1;
    }
    lds_order();  // table cleared
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      if (x[j] < 0) continue;
      u32 h = hash32((u32)x[j]) & mask;
      while (true) {
        const i32 old = atomicCAS(&K[h], -1, x[j]);
        if (old == -1 || old == x[j]) break;
        h = (h + 1) & mask;
      }
      atomicMin(&P[h], j * WAVE + lane);
      atomicAdd(&W[h], w[j]);
    }
    lds_order();  // every entry inserted
    u64* Rc = R + eoff[c];
    i32 kept = 0;
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      const i64 p = j * WAVE + lane;
      if (p >= d) continue;
      u64 word = (u64)RL_NONE << 32;
      if (x[j] >= 0) {
        u32 h = hash32((u32)x[j]) & mask;
        while (K[h] != x[j]) h = (h + 1) & mask;
        if (P[h] == (i32)p) {
          word = ((u64)(u32)x[j] << 32) | (u64)(u32)W[h];
          kept++;
        }
      }
      Rc[p] = word;
    }
    kept = wave_sum(kept);
    if (lane == 0) cnt[c] = kept;
  }
}
__global__ __launch_bounds__(256) void k_rep_medw1(const i32* lst, i64 nl, const i64* dsz,
                                                   const i64* eoff, const u64* R, const i64* ox,
                                                   i32* oa, i32* ow) {
  const int lane = lane_id();
  for (i64 e = gwave(); e < nl; e += nwaves()) {
    const i64 c = lst[e];
    const i64 d = dsz[c];
    const u64* Rc = R + eoff[c];
    u64 rv[REP_MEDW_PER];
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      const i64 p = j * WAVE + lane;
      rv[j] = p < d ? Rc[p] : ((u64)RL_NONE << 32);
    }
    i64 run = ox[c];
#pragma unroll
    for (int j = 0; j < REP_MEDW_PER; j++) {
      const bool keep = (u32)(rv[j] >> 32) != RL_NONE;
      const u64 bm = __ballot(keep);
      if (keep) {
        const i64 o = run + __popcll(bm & lanes_below(lane));
        oa[o] = (i32)(rv[j] >> 32);
        ow[o] = (i32)(u32)rv[j];
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

// Arcs v -> u with 0 <= u < n and u != v over the same row ranges as k_pairs.
// On a symmetric level whose rows hold no repeats every such pair {u, v}
// appears exactly twice, so the unique pairs are half this count.
__global__ __launch_bounds__(256) void k_count_pair_arcs(DGraph g, u64* count) {
  u64 local = 0;
  const int lane = lane_id();
  for (i64 v = gwave(); v < g.n; v += nwaves()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : g.m;
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      local += (u >= 0 && u < g.n && u != v) ? 1 : 0;
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
// cand[v] of a vertex that is not a move candidate (gains of candidates are finite)
constexpr float CAND_NO = __builtin_nanf("");

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
__global__ void k_bg_table(const i32* pc, i32 k, i64 n, int phase, float* bg, u8* src_big,
                           i32* zero_k, i32 nzero, i32* zero_2, i32* one1) {
  const i64 T = n / k;
  const float Tf = (float)T;
  if (zero_2 && gtid() < 2) zero_2[gtid()] = 0;
  if (one1 && gtid() == 0) *one1 = 1;
  if (zero_k)
    for (i64 q = gtid(); q < nzero; q += gstride()) zero_k[q] = 0;
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
                                               float* gain, i32* target, float* cand, int phase,
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
      cand[v] = (phase ? (best > -50.0f) : (best > -10.0f)) ? best : CAND_NO;  // gpu_part.cu:308-312
    }
    __syncthreads();
  }
}

// One thread per vertex whose range has at most GAIN_LIGHT arcs: the tallies
// of up to GAIN_SLOTS distinct parts live in registers and are accumulated
// in row order (exactly the reference's fp32 sums).  Vertices with more
// distinct parts are deferred to the wave kernel.
constexpr int GAIN_SLOTS = 16;
constexpr int GAIN_LIGHT = 64;


// Rows longer than GAIN_HUGE arcs: one block of 1024 threads per vertex with
// four independent loads per thread in flight; block-wide LDS tally and pick.
constexpr int GAIN_HUGE = 4096;
constexpr int HUGE_BS = 1024;


// Target and gain of every long row from its accumulated tally (a row whose
// sum of |w| exceeds 2^24 is re-tallied in fp32 in row order by one thread, as
// in the other kernels); the accumulators are cleared for the next pass.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_gain_huge_final(DGraph g, i64 nedges, const i32* part,
                                                         i32 k, const float* bg_g,
                                                         const u8* src_big, float* gain,
                                                         i32* target, float* cand, int phase,
                                                         const i32* huge, i64 nhuge,
                                                         i32* gtally, i64* gabs,
                                                         TallyRef Tout) {
  extern __shared__ i32 lds[];
  __shared__ float s_g[256 / WAVE];
  __shared__ i32 s_q[256 / WAVE];
  float* sbg = (float*)lds;
  float* ftally = (float*)(lds + 2 * k);
  for (int q = threadIdx.x; q < 2 * k; q += blockDim.x) sbg[q] = bg_g[q];
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (i64 h = blockIdx.x; h < nhuge; h += gridDim.x) {
    const i64 v = huge[h];
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const i32 cur = part[v];
    i32* gt = gtally + h * k;
    if (cur < k && cur >= 0) {
      const bool exact_int = UNIT ? (s1 - s0) <= (i64)(1 << 24) : gabs[h] <= (i64)(1 << 24);
      if (!exact_int) {
        for (int q = threadIdx.x; q < k; q += blockDim.x) ftally[q] = 0.0f;
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
      const float ccur = exact_int ? (float)gt[cur] : ftally[cur];
      const float* bgr = sbg + (src_big[cur] ? k : 0);
      float bgv = -INFINITY;
      i32 bq = INT_MAX;
      for (int q = threadIdx.x; q < k; q += blockDim.x) {
        if (q == cur) continue;
        const float cq = exact_int ? (float)gt[q] : ftally[q];
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
        for (int w2 = 1; w2 < (int)(blockDim.x / WAVE); w2++) {
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
        cand[v] = (phase ? (best > -50.0f) : (best > -10.0f)) ? best : CAND_NO;
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < k; q += blockDim.x) {
      if (Tout.t32) tally_put(Tout, v, k, q, gt[q]);
      gt[q] = 0;
    }
    if (threadIdx.x == 0) gabs[h] = 0;
    __syncthreads();
  }
}

// (s0, s1) of every long row, with the last row ending at nedges
__global__ void k_huge_bounds(DGraph g, i64 nedges, const i32* huge, i64 nhuge, i64* hb) {
  for (i64 h = gtid(); h < nhuge; h += gstride()) {
    const i64 v = huge[h];
    hb[2 * h] = g.xadj[v];
    hb[2 * h + 1] = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
  }
}

// ---- fused gain pass ---------------------------------------------------------
constexpr unsigned GAIN_FUSED_BLOCKS = 2048;  // 8 blocks of 4 waves per CU
constexpr int GAIN_CTRS = 8;                  // work counters, 128 bytes apart
// One persistent launch serves every row class.  Waves take work items from
// one counter, largest first: chunks of long rows (one wave per chunk), then
// medium rows (one wave per vertex), mid rows (4 vertices per wave, 16 lanes
// each), light rows (8 vertices per wave, 8 lanes each).  All classes run at
// once and the dynamic order balances the power-law row lengths; the next
// item's counter increment is issued before the current item is processed.
// Each wave owns 8 tallies of k ints in LDS (odd stride), zero between items.

// One vertex per G-lane group: tally, exactness, pick (the body of the
// per-class kernels).  `tally` is the group's LDS tally, zero on entry and exit.
template <int G, bool UNIT, typename PT>
__device__ __forceinline__ void gain_vertex(const DGraph& g, i64 nedges, const PT* part, i32 k,
                                            const float* sbg, const u8* src_big, float* gain,
                                            i32* target, float* cand, int phase, int all_exact,
                                            i64 v, i32* tally, u64& cutv, int gl,
                                            const TallyRef& Tout) {
  float* ftally = (float*)tally;
  const i64 s0 = g.xadj[v];
  const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
  const i32 cur = part[v];
  if (cur >= k || cur < 0) return;  // gpu_part.cu:262
  i64 absw = 0;
  i32 own = 0;  // arcs into cur are summed in a register (most arcs; one contended LDS word)
  constexpr int GU = 8;
  // software pipeline: the next step's arcs are loaded before this step's
  // gathers and tallies, so two steps of loads are in flight
  i32 uu[GU], ww[GU];
#pragma unroll
  for (int j = 0; j < GU; j++) {
    const i64 i = s0 + j * G + gl;
    uu[j] = i < s1 ? g.adj[i] : INT_MAX;
    ww[j] = (UNIT || i >= s1) ? 1 : g.w[i];
  }
  for (i64 base = s0; base < s1; base += G * GU) {
    i32 pp[GU];
#pragma unroll
    for (int j = 0; j < GU; j++) pp[j] = uu[j] < g.n ? part[uu[j]] : -1;
    const i32 cw[GU] = {ww[0], ww[1], ww[2], ww[3], ww[4], ww[5], ww[6], ww[7]};
    const i64 nb = base + G * GU;
    if (nb < s1) {
#pragma unroll
      for (int j = 0; j < GU; j++) {
        const i64 i = nb + j * G + gl;
        uu[j] = i < s1 ? g.adj[i] : INT_MAX;
        ww[j] = (UNIT || i >= s1) ? 1 : g.w[i];
      }
    }
#pragma unroll
    for (int j = 0; j < GU; j++) {
      if (pp[j] < k && pp[j] >= 0) {
        if (!UNIT && !all_exact) absw += cw[j] < 0 ? -(i64)cw[j] : (i64)cw[j];
        if (pp[j] == cur) {
          own = (i32)((u32)own + (u32)cw[j]);
        } else {
          atomicAdd(&tally[pp[j]], cw[j]);
          cutv += (u64)(i64)cw[j];
        }
      }
    }
  }
  if (!UNIT && !all_exact) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) absw += __shfl_xor(absw, o, G);
  }
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) own = (i32)((u32)own + (u32)__shfl_xor(own, o, G));
  __builtin_amdgcn_wave_barrier();
  const bool exact_int =
      UNIT ? (s1 - s0) <= (i64)(1 << 24) : (all_exact || absw <= (i64)(1 << 24));
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
  const float ccur = exact_int ? (float)own : ftally[cur];
  __builtin_amdgcn_wave_barrier();
  const float* bgr = sbg + (src_big[cur] ? k : 0);
  float bgv = -INFINITY;
  i32 bq = INT_MAX;
  for (int q = gl; q < k; q += G) {
    const i32 tq = tally[q];
    const float cq = exact_int ? (float)tq : ftally[q];
    tally[q] = 0;
    if (Tout.t32) tally_put(Tout, v, k, q, q == cur ? own : tq);  // integers: exact
    if (q == cur) continue;
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
    cand[v] = (phase ? (best > -50.0f) : (best > -10.0f)) ? best : CAND_NO;
  }
  __builtin_amdgcn_wave_barrier();
}

// One chunk of a long row by one wave: tally in the wave's LDS, merged into the
// row's global tally with integer atomics (order-free); the tally is re-zeroed.
template <bool UNIT, typename PT>
__device__ __forceinline__ void gain_chunk_wave(const DGraph& g, const PT* part, i32 k,
                                                const GainLists& L, i64 c, i32* tally,
                                                u64& cutv, int lane, int all_exact) {
  const i32 h = L.ch_row[c];
  const i32 cur = part[L.huge[h]];
  if (cur >= k || cur < 0) return;  // gpu_part.cu:262
  const i64 e0 = L.ch_beg[c], e1 = L.ch_end[c];
  i64 absw = 0;
  i32 own = 0;
  for (i64 base = e0; base < e1; base += 8 * WAVE) {
    i32 uu[8], ww[8], pp[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const i64 i = base + j * WAVE + lane;
      uu[j] = i < e1 ? g.adj[i] : INT_MAX;
      ww[j] = (UNIT || i >= e1) ? 1 : g.w[i];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) pp[j] = uu[j] < g.n ? part[uu[j]] : -1;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (pp[j] >= 0 && pp[j] < k) {
        if (!UNIT && !all_exact) absw += ww[j] < 0 ? -(i64)ww[j] : (i64)ww[j];
        if (pp[j] == cur) {
          own = (i32)((u32)own + (u32)ww[j]);
        } else {
          atomicAdd(&tally[pp[j]], ww[j]);
          cutv += (u64)(i64)ww[j];
        }
      }
    }
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) own = (i32)((u32)own + (u32)__shfl_xor(own, o));
  if (lane == 0 && own) atomicAdd(&L.gtally[(i64)h * k + cur], own);
  __builtin_amdgcn_wave_barrier();
  for (int q = lane; q < k; q += WAVE) {
    const i32 tq = tally[q];
    tally[q] = 0;
    if (tq) atomicAdd(&L.gtally[(i64)h * k + q], tq);
  }
  if (!UNIT && !all_exact) {
    absw = wave_sum(absw);
    if (lane == 0 && absw) atomicAdd((unsigned long long*)&L.gabs[h], (unsigned long long)absw);
  }
  __builtin_amdgcn_wave_barrier();
}

template <bool UNIT, typename PT>
__global__ __launch_bounds__(256) void k_gain_fused(DGraph g, i64 nedges, const PT* part, i32 k,
                                                    const float* bg_g, const u8* src_big,
                                                    float* gain, i32* target, float* cand,
                                                    int phase, GainLists L, u32* ctr,
                                                    u64* cut_acc, TallyRef Tout) {
  extern __shared__ i32 lds[];
  float* sbg = (float*)lds;
  for (int q = threadIdx.x; q < 2 * k; q += blockDim.x) sbg[q] = bg_g[q];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int ks = k | 1;
  i32* wt = lds + 2 * k + wv * 8 * ks;
  for (int q = lane; q < 8 * ks; q += WAVE) wt[q] = 0;
  __syncthreads();
  const int all_exact = L.all_exact;
  u64 cutv = 0;
  // items: long-row chunks, medium rows, 16 mid rows, 64 vertices (light rows)
  const i64 e1 = L.nch, e2 = e1 + L.nmedium, e3 = e2 + (L.nmid + 15) / 16,
            e4 = e3 + (g.n + 63) / 64;
  // GAIN_CTRS counters, one per 128-byte line, hand out interleaved items
  const int cq = (blockIdx.x * 4 + wv) % GAIN_CTRS;
  u32* myctr = ctr + cq * 32;
  u32 it = 0;
  if (lane == 0) it = atomicAdd(myctr, 1u);
  i64 item = (i64)__shfl(it, 0) * GAIN_CTRS + cq;
  while (item < e4) {
    u32 nxt = 0;
    if (lane == 0) nxt = atomicAdd(myctr, 1u);
    if (item < e1) {
      gain_chunk_wave<UNIT, PT>(g, part, k, L, item, wt, cutv, lane, all_exact);
    } else if (item < e2) {
      gain_vertex<64, UNIT, PT>(g, nedges, part, k, sbg, src_big, gain, target, cand, phase,
                                all_exact, L.medium[item - e1], wt, cutv, lane, Tout);
    } else if (item < e3) {
      for (int r = 0; r < 4; r++) {
        const i64 idx = (item - e2) * 16 + r * 4 + (lane >> 4);
        if (idx < L.nmid)
          gain_vertex<16, UNIT, PT>(g, nedges, part, k, sbg, src_big, gain, target, cand, phase,
                                    all_exact, L.mid[idx], wt + (lane >> 4) * ks, cutv,
                                    lane & 15, Tout);
      }
    } else {
      for (int r = 0; r < 8; r++) {
        const i64 v = (item - e3) * 64 + r * 8 + (lane >> 3);
        bool light = v < g.n;
        if (light) {
          const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
          light = s1 - g.xadj[v] <= GAIN_LIGHT;
        }
        if (light)
          gain_vertex<8, UNIT, PT>(g, nedges, part, k, sbg, src_big, gain, target, cand, phase,
                                   all_exact, v, wt + (lane >> 3) * ks, cutv, lane & 7, Tout);
      }
    }
    item = (i64)__shfl(nxt, 0) * GAIN_CTRS + cq;
  }
  if (cut_acc) block_add_u64(cutv, cut_acc);
}

// ---- refinement from kept tallies --------------------------------------------
// On a level whose rows are simple and weight-symmetric, T[v][q] (the weight of
// v's arcs into part q, the tally the gain kernels gather) is kept across
// passes: only the neighbours of the moved vertices change, by -w on the old
// part and +w on the new one, and a pass reads T instead of gathering part[u]
// over every arc.  Integer tallies make the two ways of computing T equal.

// The pick of gain_vertex from T, 16 lanes per vertex; also the vertex's
// share of the edge cut (its arcs into other parts).
// Levels whose tallies are all int rows (row v at v * tally_stride(k)): G lanes
// per vertex, R 16-byte loads per lane in flight.
template <int G, int R>
__global__ __launch_bounds__(256) void k_gain_tallies(i64 n, const i32* part, TallyRef T,
                                                      i32 k, const float* bg_g,
                                                      const u8* src_big, float* gain,
                                                      i32* target, float* cand, int phase,
                                                      u64* cut_acc) {
  extern __shared__ float sbgt[];
  for (int q = threadIdx.x; q < 2 * k; q += blockDim.x) sbgt[q] = bg_g[q];
  __syncthreads();
  const int gl = threadIdx.x & (G - 1);
  const i64 ngroups = (i64)gridDim.x * (blockDim.x / G);
  const i64 ts = tally_stride(k);
  const int nv4 = (int)(ts >> 2);
  u64 cutv = 0;
  for (i64 v = (i64)blockIdx.x * (blockDim.x / G) + threadIdx.x / G; v < n; v += ngroups) {
    const i32 cur = part[v];
    if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
    const i32* Tv = T.t32 + v * ts;
    const int4* T4 = (const int4*)Tv;  // rows are 16-byte aligned (stride multiple of 4)
    const float ccur = (float)Tv[cur];
    const float* bgr = sbgt + (src_big[cur] ? k : 0);
    float bgv = -INFINITY;
    i32 bq = INT_MAX;
    i64 other = 0;
    for (int c0 = 0; c0 < nv4; c0 += G * R) {
      int4 t4r[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int c = c0 + gl + r * G;
        t4r[r] = c < nv4 ? T4[c] : make_int4(0, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int c = c0 + gl + r * G;
        const i32 tv4[4] = {t4r[r].x, t4r[r].y, t4r[r].z, t4r[r].w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const i32 q = 4 * c + j;
          if (q >= k || q == cur) continue;
          const i32 tq = tv4[j];
          other += tq;
          const float eg = (float)tq - ccur;
          const float gg = eg + bgr[q];
          if (gg > bgv || (gg == bgv && q < bq)) {
            bgv = gg;
            bq = q;
          }
        }
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
    cutv += (u64)other;
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
      cand[v] = (phase ? (best > -50.0f) : (best > -10.0f)) ? best : CAND_NO;
    }
  }
  if (cut_acc) block_add_u64(cutv, cut_acc);
}

// k <= 64 on a level with byte rows: one lane per vertex.  The first nb_byte
// blocks take the byte rows (up to four 16-byte loads each, issued together
// with the other 63 lanes' rows), the others the int rows from their list.
// The pick runs over q in lockstep across the wave without branches: the
// balance terms come from LDS four at a time (entries past k hold NaN, which
// never wins), and since q rises, the first q of the largest gain is the one a
// strict comparison keeps.  The arcs into other parts are the row's sum (its
// padding holds zero) less the own tally.
constexpr int GT1_MAXK = 64;
__device__ __forceinline__ void gt1_pick(i32 q, float tq, i32 cur, float ccur, float b,
                                         float& bgv, i32& bq) {
  const float gg = (tq - ccur) + b;
  const bool better = gg > bgv && q != cur;
  bgv = better ? gg : bgv;
  bq = better ? q : bq;
}
__device__ __forceinline__ void gt1_store(i64 v, i32 cur, float bgv, i32 bq, float* gain,
                                          i32* target, float* cand, int phase) {
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
  cand[v] = (phase ? (best > -50.0f) : (best > -10.0f)) ? best : CAND_NO;
}
__global__ __launch_bounds__(256) void k_gain_tallies1(i64 n, const i32* part, TallyRef T,
                                                       i32 k, const float* bg_g,
                                                       const u8* src_big, float* gain,
                                                       i32* target, float* cand, int phase,
                                                       u64* cut_acc, unsigned nb_byte) {
  // two rows of GT1_MAXK balance terms (source part not big / big)
  __shared__ __align__(16) float sbgt[2 * GT1_MAXK];
  for (int i = threadIdx.x; i < 2 * GT1_MAXK; i += blockDim.x) {
    const int q = i & (GT1_MAXK - 1);
    sbgt[i] = q < k ? bg_g[(i >= GT1_MAXK ? k : 0) + q] : CAND_NO;
  }
  __syncthreads();
  const int nc = (k + 15) >> 4;  // 16-tally chunks
  u64 cutv = 0;
  if (blockIdx.x < nb_byte) {
    const i64 ts8 = tally8_stride_d(k);
    for (i64 v0 = (i64)blockIdx.x * blockDim.x; v0 < n; v0 += (i64)nb_byte * blockDim.x) {
      const i64 v = v0 + threadIdx.x;
      if (v >= n || ((T.bigmask[v >> 5] >> (v & 31)) & 1u)) continue;
      const i32 cur = part[v];
      if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
      const float4* B4 = (const float4*)(sbgt + (src_big[cur] ? GT1_MAXK : 0));
      const int4* R16 = (const int4*)(T.t8 + v * ts8);
      int4 c4[4];
#pragma unroll
      for (int c = 0; c < 4; c++) c4[c] = c < nc ? R16[c] : make_int4(0, 0, 0, 0);
      const u32 own = T.t8[v * ts8 + cur];
      const float ccur = (float)own;
      float bgv = -INFINITY;
      i32 bq = INT_MAX;
      u32 tot = 0;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        if (c >= nc) break;
        const u32 w4[4] = {(u32)c4[c].x, (u32)c4[c].y, (u32)c4[c].z, (u32)c4[c].w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int q0 = 16 * c + 4 * i;
          const u32 w = w4[i];
          tot = __builtin_amdgcn_sad_u8(w, 0u, tot);
          const float4 b = B4[q0 >> 2];
          gt1_pick(q0, (float)(w & 255u), cur, ccur, b.x, bgv, bq);
          gt1_pick(q0 + 1, (float)((w >> 8) & 255u), cur, ccur, b.y, bgv, bq);
          gt1_pick(q0 + 2, (float)((w >> 16) & 255u), cur, ccur, b.z, bgv, bq);
          gt1_pick(q0 + 3, (float)(w >> 24), cur, ccur, b.w, bgv, bq);
        }
      }
      cutv += (u64)(tot - own);
      gt1_store(v, cur, bgv, bq, gain, target, cand, phase);
    }
  } else {
    const i64 ts = tally_stride(k);
    const int nv4 = (int)(ts >> 2);
    const i64 stride = (i64)(gridDim.x - nb_byte) * blockDim.x;
    for (i64 r = (i64)(blockIdx.x - nb_byte) * blockDim.x + threadIdx.x; r < T.nbig;
         r += stride) {
      const i64 v = T.bigv[r];
      const i32 cur = part[v];
      if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
      const float4* B4 = (const float4*)(sbgt + (src_big[cur] ? GT1_MAXK : 0));
      const int4* R4 = (const int4*)(T.t32 + r * ts);  // bix[v] == r
      const i32 own = T.t32[r * ts + cur];
      const float ccur = (float)own;
      float bgv = -INFINITY;
      i32 bq = INT_MAX;
      i64 tot = 0;
#pragma unroll 1
      for (int c = 0; c < nc; c++) {
        int4 c4[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
          c4[i] = 4 * c + i < nv4 ? R4[4 * c + i] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int q0 = 16 * c + 4 * i;
          const float4 b = B4[q0 >> 2];
          tot += (i64)c4[i].x + c4[i].y + c4[i].z + c4[i].w;
          gt1_pick(q0, (float)c4[i].x, cur, ccur, b.x, bgv, bq);
          gt1_pick(q0 + 1, (float)c4[i].y, cur, ccur, b.y, bgv, bq);
          gt1_pick(q0 + 2, (float)c4[i].z, cur, ccur, b.z, bgv, bq);
          gt1_pick(q0 + 3, (float)c4[i].w, cur, ccur, b.w, bgv, bq);
        }
      }
      cutv += (u64)(tot - own);
      gt1_store(v, cur, bgv, bq, gain, target, cand, phase);
    }
  }
  if (cut_acc) block_add_u64(cutv, cut_acc);
}

// 64 < k <= TALLY16_MAXK on a level with short rows: G lanes per vertex, each
// reading every G-th 16-byte chunk (8 tallies) of the row with all its loads
// in flight; the pick is gt1's (NaN-padded balance rows in LDS, strict first
// maximum within the lane), then the lanes' picks combine by (gain, lower q).
// The blocks past nb_short take the int rows from their list, one lane each.
template <int G>
__global__ __launch_bounds__(256) void k_gain_tallies16(i64 n, const i32* part, TallyRef T,
                                                        i32 k, const float* bg_g,
                                                        const u8* src_big, float* gain,
                                                        i32* target, float* cand, int phase,
                                                        u64* cut_acc, unsigned nb_short) {
  constexpr int KM = TALLY16_MAXK;
  __shared__ __align__(16) float sbgt[2 * KM];
  for (int i = threadIdx.x; i < 2 * KM; i += blockDim.x) {
    const int q = i & (KM - 1);
    sbgt[i] = q < k ? bg_g[(i >= KM ? k : 0) + q] : CAND_NO;
  }
  __syncthreads();
  u64 cutv = 0;
  if (blockIdx.x < nb_short) {
    const i64 kp = tally16_stride_d(k);
    const int nch = (int)(kp >> 3);  // 16-byte chunks of a row
    constexpr int MAXC = KM / 8 / G;  // chunks per lane at most
    const int gl = threadIdx.x & (G - 1);
    const i64 ngroups = (i64)nb_short * (blockDim.x / G);
    for (i64 v = (i64)blockIdx.x * (blockDim.x / G) + threadIdx.x / G; v < n; v += ngroups) {
      if ((T.bigmask[v >> 5] >> (v & 31)) & 1u) continue;
      const i32 cur = part[v];
      if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
      const float4* B4 = (const float4*)(sbgt + (src_big[cur] ? KM : 0));
      const int4* R = (const int4*)(T.t16 + v * kp);
      int4 c4[MAXC];
#pragma unroll
      for (int j = 0; j < MAXC; j++) {
        const int c = gl + j * G;
        c4[j] = c < nch ? R[c] : make_int4(0, 0, 0, 0);
      }
      const u32 own = T.t16[v * kp + cur];
      const float ccur = (float)own;
      float bgv = -INFINITY;
      i32 bq = INT_MAX;
      u32 tot = 0;
#pragma unroll
      for (int j = 0; j < MAXC; j++) {
        const int c = gl + j * G;
        if (c >= nch) break;
        const u32 w4[4] = {(u32)c4[j].x, (u32)c4[j].y, (u32)c4[j].z, (u32)c4[j].w};
        const float4 b0 = B4[2 * c], b1 = B4[2 * c + 1];
        const float bb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int h = 0; h < 8; h++) {
          const u32 t = (w4[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
          tot += t;
          gt1_pick(8 * c + h, (float)t, cur, ccur, bb[h], bgv, bq);
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
        tot += __shfl_xor(tot, o, G);
      }
      if (gl == 0) {
        cutv += (u64)(tot - own);
        gt1_store(v, cur, bgv, bq, gain, target, cand, phase);
      }
    }
  } else {
    const i64 ts = tally_stride(k);
    const int nv4 = (int)(ts >> 2);
    const int nc = (k + 15) >> 4;
    const i64 stride = (i64)(gridDim.x - nb_short) * blockDim.x;
    for (i64 r = (i64)(blockIdx.x - nb_short) * blockDim.x + threadIdx.x; r < T.nbig;
         r += stride) {
      const i64 v = T.bigv[r];
      const i32 cur = part[v];
      if (cur >= k || cur < 0) continue;  // gpu_part.cu:262
      const float4* B4 = (const float4*)(sbgt + (src_big[cur] ? KM : 0));
      const int4* R4 = (const int4*)(T.t32 + r * ts);  // bix[v] == r
      const i32 own = T.t32[r * ts + cur];
      const float ccur = (float)own;
      float bgv = -INFINITY;
      i32 bq = INT_MAX;
      i64 tot = 0;
#pragma unroll 1
      for (int c = 0; c < nc; c++) {
        int4 c4[4];
#pragma unroll
        for (int i = 0; i < 4; i++)
          c4[i] = 4 * c + i < nv4 ? R4[4 * c + i] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int q0 = 16 * c + 4 * i;
          const float4 b = B4[q0 >> 2];
          tot += (i64)c4[i].x + c4[i].y + c4[i].z + c4[i].w;
          gt1_pick(q0, (float)c4[i].x, cur, ccur, b.x, bgv, bq);
          gt1_pick(q0 + 1, (float)c4[i].y, cur, ccur, b.y, bgv, bq);
          gt1_pick(q0 + 2, (float)c4[i].z, cur, ccur, b.z, bgv, bq);
          gt1_pick(q0 + 3, (float)c4[i].w, cur, ccur, b.w, bgv, bq);
        }
      }
      cutv += (u64)(tot - own);
      gt1_store(v, cur, bgv, bq, gain, target, cand, phase);
    }
  }
  if (cut_acc) block_add_u64(cutv, cut_acc);
}

__global__ void k_big_list(const i32* bix, i64 n, i32* bigv) {
  for (i64 v = gtid(); v < n; v += gstride()) {
    const i32 b = bix[v];
    if (b >= 0) bigv[b] = (i32)v;
  }
}

// Edge cut from kept tallies (gpu_part.cu:400-422 on a level whose tallies
// T[v][q] are the weights of v's arcs, over its row range, into part q):
// sum over v with part[v] in [0, k) of sum_{q != part[v]} T[v][q], into the
// accumulator slots.  Four lanes per vertex, 16-byte loads.
__global__ __launch_bounds__(256) void k_tally_cut(i64 n, const i32* part, TallyRef T, i32 k,
                                                   u64* cut_acc) {
  const int gl = threadIdx.x & 3;
  const i64 ngroups = (i64)gridDim.x * (blockDim.x / 4);
  const i64 ts = tally_stride(k), ts8 = tally8_stride_d(k);
  const int nv4 = (int)(ts >> 2), nv16 = (int)(ts8 >> 4);
  u64 cutv = 0;
  for (i64 v = (i64)blockIdx.x * (blockDim.x / 4) + threadIdx.x / 4; v < n; v += ngroups) {
    const i32 cur = part[v];
    if (cur >= k || cur < 0) continue;
    const i64 b = tally_int_row(T, v);
    i64 other = 0;
    if (b >= 0) {
      const int4* T4 = (const int4*)(T.t32 + (i64)b * ts);
      for (int c = gl; c < nv4; c += 4) {
        const int4 t = T4[c];
        const i32 tv[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const i32 q = 4 * c + j;
          if (q < k && q != cur) other += tv[j];
        }
      }
    } else if (T.t16) {
      const int4* S4 = (const int4*)(T.t16 + v * tally16_stride_d(k));
      const int nc8 = (int)(tally16_stride_d(k) >> 3);
      for (int c = gl; c < nc8; c += 4) {
        const int4 t = S4[c];
        const u32 w4[4] = {(u32)t.x, (u32)t.y, (u32)t.z, (u32)t.w};
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const i32 q = 8 * c + j;
          if (q < k && q != cur) other += (w4[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
        }
      }
    } else {
      const int4* T16 = (const int4*)(T.t8 + v * ts8);
      for (int c = gl; c < nv16; c += 4) {
        const int4 t = T16[c];
        const u32 w4[4] = {(u32)t.x, (u32)t.y, (u32)t.z, (u32)t.w};
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const i32 q = 16 * c + j;
          if (q < k && q != cur) other += (w4[j >> 2] >> (8 * (j & 3))) & 255u;
        }
      }
    }
    cutv += (u64)other;
  }
  block_add_u64(cutv, cut_acc);
}

// After a commit: every applied move (v: src -> dst) shifts the weight of each
// of v's arcs (v, u) from T[u][src] to T[u][dst]; the reverse arc (u, v) has
// the same weight on these levels.  T[n-1] is skipped when the last row's
// range is not the whole row (it is rebuilt by k_tally_last then).
__global__ __launch_bounds__(256) void k_tally_update(DGraph g, TallyRef T, i32 k, const i32* app,
                                                      const i32* napp, const i32* mv_v,
                                                      const i32* mv_src, const i32* mv_dst,
                                                      int unit, int skip_last) {
  const int lane = lane_id();
  const i64 cnt = *napp;
  for (i64 e = gwave(); e < cnt; e += nwaves()) {
    const i32 id = app[e];
    const i32 v = mv_v[id], src = mv_src[id], dst = mv_dst[id];
    const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
    for (i64 i = s0 + lane; i < s1; i += WAVE) {
      const i32 u = g.adj[i];
      if (u < 0 || u >= g.n || (skip_last && u == g.n - 1)) continue;
      const i32 w = unit ? 1 : g.w[i];
      tally_move(T, u, k, src, dst, w);
    }
  }
}

// Rows kept as bytes (limit 255) or shorts (limit 65535): every arc weighs >= 0
// and the row (the last row ending at nedges) sums to at most the limit; the
// last row always keeps ints (its range may change with nedges).
__global__ __launch_bounds__(256) void k_tally_class(DGraph g, i64 nedges, int unit, int all_int,
                                                     i64 limit, i32* bix, u32* bigmask,
                                                     i32* nbig) {
  for (i64 v0 = (i64)blockIdx.x * blockDim.x; v0 < g.n; v0 += (i64)gridDim.x * blockDim.x) {
    const i64 v = v0 + threadIdx.x;
    bool big = true;
    if (v < g.n && !all_int && v != g.n - 1) {
      const i64 s0 = g.xadj[v], s1 = g.xadj[v + 1];
      if (unit) {
        big = s1 - s0 > limit;
      } else {
        big = false;
        i64 sum = 0;
        for (i64 i = s0; i < s1; i++) {
          const i32 w = g.w[i];
          sum += w;
          if (w < 0 || sum > limit) {
            big = true;
            break;
          }
        }
      }
    }
    const bool want = v < g.n && big;
    const u64 m = __ballot(want);
    i32 base = 0;
    const int lane = lane_id();
    const int leader = m ? __ffsll((unsigned long long)m) - 1 : 0;
    if (m && lane == leader) base = atomicAdd(nbig, (i32)__popcll(m));
    base = __shfl(base, leader);
    if (v < g.n) bix[v] = want ? base + (i32)__popcll(m & lanes_below(lane)) : -1;
    // 64 consecutive vertices per wave (v0 and blockDim are multiples of 64)
    if (lane < 2 && v - lane + 32 * lane < g.n) bigmask[(v - lane) / 32 + lane] = (u32)(m >> (32 * lane));
  }
}

// T[n-1] from its range [xadj[n-1], nedges) (one block; the last row keeps ints)
__global__ __launch_bounds__(256) void k_tally_last(DGraph g, i64 nedges, const i32* part,
                                                    TallyRef T, i32 k, int unit) {
  const i64 v = g.n - 1;
  i32* Tv = T.t32 + tally_int_row(T, v) * tally_stride(k);
  for (int q = threadIdx.x; q < k; q += blockDim.x) Tv[q] = 0;
  __syncthreads();
  for (i64 i = g.xadj[v] + threadIdx.x; i < nedges; i += blockDim.x) {
    const i32 u = g.adj[i];
    if (u >= g.n) continue;
    const i32 p = part[u];
    if (p >= 0 && p < k) atomicAdd(&Tv[p], unit ? 1 : g.w[i]);
  }
}

// Heavy rows split into medium (64 < len <= 4096, one wave) and huge (> 4096,
// one block): entries of `heavy` are partitioned in place by the census.
// *over = 1 when some row's sum of |w| (over [s0, s1), the last row ending at
// nedges) exceeds 2^24; rows longer than heavy_deg are summed by one wave
// each from the heavy list.
// A row whose degree times the level's largest |w| is at most 2^24 (and that
// lies within the level's arcs, which that maximum covers) cannot pass 2^24
// and is not summed.
__device__ __forceinline__ bool row_abs_bounded(const DGraph& g, i64 s1, i64 d, u64 maxabs) {
  return s1 <= g.m && (u64)d * maxabs <= (u64)(1 << 24);
}
// over[0] counts the rows whose sum of |w| passes 2^24 and lists them in
// over + 1 (in no particular order).
__global__ void k_row_abs_light(DGraph g, i64 nedges, i64 heavy_deg, const u64* maxabs_p,
                                i32* over) {
  const u64 maxabs = *maxabs_p;
  for (i64 v = gtid(); v < g.n; v += gstride()) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > heavy_deg || row_abs_bounded(g, s1, s1 - s0, maxabs)) continue;
    i64 a = 0;
    for (i64 i = s0; i < s1; i++) a += g.w[i] < 0 ? -(i64)g.w[i] : (i64)g.w[i];
    if (a > (i64)(1 << 24)) over[1 + atomicAdd(&over[0], 1)] = (i32)v;
  }
}
__global__ __launch_bounds__(256) void k_row_abs_heavy(DGraph g, i64 nedges, const i32* list,
                                                       i64 cnt, const u64* maxabs_p, i32* over) {
  const int lane = lane_id();
  const u64 maxabs = *maxabs_p;
  for (i64 e = gwave(); e < cnt; e += nwaves()) {
    const i64 v = list[e];
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (row_abs_bounded(g, s1, s1 - s0, maxabs)) continue;
    i64 a = 0;
    for (i64 i = s0 + lane; i < s1; i += WAVE) a += g.w[i] < 0 ? -(i64)g.w[i] : (i64)g.w[i];
    a = wave_sum(a);
    if (lane == 0 && a > (i64)(1 << 24)) over[1 + atomicAdd(&over[0], 1)] = (i32)v;
  }
}

constexpr int GAIN_MID = 512;
__global__ void k_split_heavy(DGraph g, i64 nedges, const i32* heavy, i64 nheavy, i32* mid,
                              i32* nmid, i32* medium, i32* nmedium, i32* huge, i32* nhuge) {
  for (i64 e = gtid(); e < nheavy; e += gstride()) {
    const i32 v = heavy[e];
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const i64 d = s1 - s0;
    wave_append(d > GAIN_HUGE, v, huge, nhuge);
    wave_append(d > GAIN_MID && d <= GAIN_HUGE, v, medium, nmedium);
    wave_append(d <= GAIN_MID, v, mid, nmid);
  }
}

// gpu_part.cu:316-333 (+ the move-eligibility test of 337-343 folded in).
__device__ __forceinline__ bool move_ok(i32 src, i32 dst, i32 k) {
  return src < k && dst < k && src != dst && src >= 0 && dst >= 0;
}

// wit[v] (optional) keeps, from an earlier pass on the level, a neighbour
// that beat v then; when it still beats v, v loses without its row being
// read.  The rule is the same ("some neighbour beats v"), so a stale witness
// only costs the row read.  The last row (whose range ends at nedges) keeps
// none.
__device__ __forceinline__ bool token_beats(float gu, i32 u, float gv, i64 v) {
  return gu > gv + 10.0f || (gu == gv && (i64)u < v);
}
__device__ __forceinline__ bool token_witness_beats(const i32* wit, const float* cand, i64 n,
                                                    i64 v, float gv) {
  if (!wit || v == n - 1) return false;
  const i32 x = wit[v];
  return x >= 0 && x < n && token_beats(cand[x], x, gv, v);
}

// cand[u] is u's gain when u is a candidate and NaN otherwise, so one load
// answers both questions and every comparison with a non-candidate is false.
// One launch serves both row classes at once: the first `bh` blocks (issued
// first, they run longest) take the heavy rows (one wave per vertex, 64
// neighbours per step: the first losing neighbour usually comes early, so
// wider steps only add traffic), the others the light rows (<= heavy_deg
// arcs).  G = 1: one thread per light vertex, 8 neighbours per step; G > 1:
// G lanes per light vertex, G * TU neighbours per step (the lanes read
// consecutive arcs, and the group stops at the first step with a loser).
template <int G, int TU>
__device__ __forceinline__ void token_light_groups(const DGraph& g, i64 nedges,
                                                   const float* cand, const i32* part,
                                                   const i32* target, i32 k, u8* flag,
                                                   i64 heavy_deg, unsigned bh) {
  const int gl = threadIdx.x & (G - 1);
  const int gbase = lane_id() & ~(G - 1);
  constexpr u64 GMASK = (G == 64) ? ~0ull : ((1ull << G) - 1);
  const i64 ngroups = (i64)(gridDim.x - bh) * (blockDim.x / G);
  for (i64 v = (i64)(blockIdx.x - bh) * (blockDim.x / G) + threadIdx.x / G; v < g.n;
       v += ngroups) {
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    if (s1 - s0 > heavy_deg) continue;
    const float gv = cand[v];
    if (!(gv == gv)) {
      if (gl == 0) flag[v] = 0;
      continue;
    }
    const float gv10 = gv + 10.0f;
    bool lose = false;
    for (i64 base = s0; base < s1; base += G * TU) {
      i32 uu[TU];
#pragma unroll
      for (int j = 0; j < TU; j++) {
        const i64 i = base + gl + (i64)j * G;
        uu[j] = i < s1 ? g.adj[i] : INT_MAX;
      }
      bool l = false;
#pragma unroll
      for (int j = 0; j < TU; j++) {
        if (uu[j] < g.n) {
          const float gu = cand[uu[j]];
          l |= gu > gv10 || (gu == gv && uu[j] < v);
        }
      }
      if ((__ballot(l) >> gbase) & GMASK) {
        lose = true;
        break;
      }
    }
    if (gl == 0) flag[v] = (!lose && move_ok(part[v], target[v], k)) ? 1 : 0;
  }
}

// G = 0: one wave per 64 consecutive vertices, whose light candidate rows'
// arcs form one stream, 64 entries per step (lane l takes the step's entry l,
// its row found by binary search over the rows' offsets in the step, held one
// per lane), so the row loads are coalesced.  The stream is rebuilt at every
// step from the rows still undecided: a row that met a losing neighbour drops
// out at once and its lanes go to the other rows' next entries.
__device__ __forceinline__ void token_light_stream(const DGraph& g, i64 nedges,
                                                   const float* cand, const i32* part,
                                                   const i32* target, i32 k, u8* flag,
                                                   i64 heavy_deg, unsigned bh, i32* wit) {
  const int lane = lane_id();
  const i64 ngw = (i64)(gridDim.x - bh) * (blockDim.x / WAVE);
  for (i64 gw = (i64)(blockIdx.x - bh) * (blockDim.x / WAVE) + (threadIdx.x >> 6);
       gw * WAVE < g.n; gw += ngw) {
    const i64 v = gw * WAVE + lane;
    const bool valid = v < g.n;
    const i64 s0 = valid ? g.xadj[v] : 0;
    const i64 s1 = valid ? ((v + 1 < g.n) ? g.xadj[v + 1] : nedges) : 0;
    const bool light = valid && s1 - s0 <= heavy_deg;
    const float gv = light ? cand[v] : __int_as_float(0x7fc00000);
    const bool isc = gv == gv;
    const float gv10 = gv + 10.0f;
    // entries of this lane's row not yet read (the last row ends at nedges,
    // which may lie before its start: none); 0 once the row has lost
    i32 rem = (isc && s1 > s0) ? (i32)(s1 - s0) : 0;
    i64 next = s0;  // the row's next entry
    bool lost = false;
    if (rem > 0 && token_witness_beats(wit, cand, g.n, v, gv)) {
      lost = true;
      rem = 0;
    }
    i32 wnew = -1;  // a neighbour found beating v in this pass
    while (true) {
      i32 incl = rem;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const i32 t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      if (__shfl(incl, WAVE - 1) == 0) break;
      const i32 start = incl - rem;  // this row's first slot in the step
      // the last row whose first slot is <= lane (rows with nothing left share
      // the next row's slot and lie before it)
      int r = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const int q = r + step;
        if (__shfl(start, q) <= lane) r = q;
      }
      const i32 rst = __shfl(start, r), rrem = __shfl(rem, r);
      const bool act = lane - rst < rrem;
      const i64 at = __shfl(next, r) + (lane - rst);
      const float rg = __shfl(gv, r), rg10 = __shfl(gv10, r);
      bool l = false;
      const i32 u = act ? g.adj[at] : -1;
      if (u >= 0 && u < g.n) {
        const float gu = cand[u];
        l = gu > rg10 || (gu == rg && (i64)u < gw * WAVE + r);
      }
      // rows with a losing neighbour in this step (and one such neighbour each)
      u64 lm = 0;
      u64 m = __ballot(l);
      while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        const int rb = __shfl(r, b);
        const i32 ub = __shfl(u, b);
        lm |= 1ull << rb;
        if (lane == rb) wnew = ub;
        m &= ~__ballot(l && r == rb);
      }
      // this row's entries read in the step: min(rem, 64 - start) when start < 64
      const i32 took = start < WAVE ? (rem < WAVE - start ? rem : WAVE - start) : 0;
      next += took;
      rem -= took;
      if ((lm >> lane) & 1) {
        lost = true;
        rem = 0;
      }
    }
    if (light) flag[v] = (isc && !lost && move_ok(part[v], target[v], k)) ? 1 : 0;
    if (wit && wnew >= 0 && v != g.n - 1) wit[v] = wnew;
  }
}

template <int G>
__global__ __launch_bounds__(256) void k_token(DGraph g, i64 nedges, const float* cand,
                                               const i32* part, const i32* target, i32 k,
                                               u8* flag, i64 heavy_deg, const i32* list,
                                               i64 cnt, unsigned bh, i32* wit) {
  if (blockIdx.x >= bh && G == 0) {
    token_light_stream(g, nedges, cand, part, target, k, flag, heavy_deg, bh, wit);
    return;
  }
  if (blockIdx.x >= bh && G > 1) {
    token_light_groups<(G > 1 ? G : 2), 4>(g, nedges, cand, part, target, k, flag, heavy_deg,
                                           bh);
    return;
  }
  if (blockIdx.x >= bh) {
    constexpr int TU = 8;
    const i64 stride = (i64)(gridDim.x - bh) * blockDim.x;
    for (i64 v = (i64)(blockIdx.x - bh) * blockDim.x + threadIdx.x; v < g.n; v += stride) {
      const i64 s0 = g.xadj[v];
      const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
      if (s1 - s0 > heavy_deg) continue;
      const float gv = cand[v];
      if (!(gv == gv)) {
        flag[v] = 0;
        continue;
      }
      const float gv10 = gv + 10.0f;
      bool lose = s1 > s0 && token_witness_beats(wit, cand, g.n, v, gv);
      i32 wnew = -1;
      for (i64 base = s0; base < s1 && !lose; base += TU) {
        i32 uu[TU];
#pragma unroll
        for (int j = 0; j < TU; j++) uu[j] = base + j < s1 ? g.adj[base + j] : INT_MAX;
#pragma unroll
        for (int j = 0; j < TU; j++) {
          if (uu[j] >= 0 && uu[j] < g.n) {
            const float gu = cand[uu[j]];
            const bool l = gu > gv10 || (gu == gv && uu[j] < v);
            if (l && !lose) wnew = uu[j];
            lose |= l;
          }
        }
      }
      flag[v] = (!lose && move_ok(part[v], target[v], k)) ? 1 : 0;
      if (wit && wnew >= 0 && v != g.n - 1) wit[v] = wnew;
    }
    return;
  }
  const int lane = lane_id();
  const i64 w0 = ((i64)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const i64 nw = ((i64)bh * blockDim.x) / WAVE;
  for (i64 e = w0; e < cnt; e += nw) {
    const i32 v = list[e];
    const float gv = cand[v];
    if (!(gv == gv)) {
      if (lane == 0) flag[v] = 0;
      continue;
    }
    const i64 s0 = g.xadj[v];
    const i64 s1 = (v + 1 < g.n) ? g.xadj[v + 1] : nedges;
    const float gv10 = gv + 10.0f;
    bool lose = s1 > s0 && token_witness_beats(wit, cand, g.n, v, gv);
    for (i64 base = s0; base < s1 && !lose; base += WAVE) {
      const i64 i = base + lane;
      bool l = false;
      i32 u = -1;
      if (i < s1) {
        u = g.adj[i];
        if (u >= 0 && u < g.n) {
          const float gu = cand[u];
          l = gu > gv10 || (gu == gv && u < v);
        }
      }
      const u64 lm = __ballot(l);
      if (lm) {
        lose = true;
        const i32 ub = __shfl(u, __ffsll((unsigned long long)lm) - 1);
        if (wit && lane == 0 && v != g.n - 1) wit[v] = ub;
      }
    }
    if (lane == 0) flag[v] = (!lose && move_ok(part[v], target[v], k)) ? 1 : 0;
  }
}

// gpu_part.cu:337-350 under SEQ: winners in ascending vertex order.  Block b
// of the compaction owns vertices [b per, (b + 1) per): k_flag_count counts
// its winners, k_flag_offsets scans the counts (one block), and
// k_scatter_moves writes the block's winners in order from its offset.
constexpr int CMP_BS = 256;
constexpr int CMP_MAXB = 4096;
__device__ __forceinline__ i32 block_sum_i32(i32 v, i32* sw) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sw[wv] = v;
  __syncthreads();
  i32 t = 0;
  for (int w2 = 0; w2 < CMP_BS / WAVE; w2++) t += sw[w2];
  __syncthreads();
  return t;
}
__global__ __launch_bounds__(CMP_BS) void k_flag_count(const u8* flag, i64 n, i64 per, i32* bcnt) {
  __shared__ i32 sw[CMP_BS / WAVE];
  const i64 b0 = (i64)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  i32 c = 0;
  for (i64 v = b0 + threadIdx.x; v < b1; v += CMP_BS) c += flag[v];
  c = block_sum_i32(c, sw);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = c;
}
// one block: boff = exclusive scan of bcnt[0..nb), *total = the sum
__global__ __launch_bounds__(1024) void k_flag_offsets(const i32* bcnt, i32 nb, i32* boff,
                                                       i64* total) {
  __shared__ i32 sw[1024 / WAVE];
  constexpr int PER = CMP_MAXB / 1024;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  i32 v[PER], s = 0;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int i = threadIdx.x * PER + r;
    v[r] = i < nb ? bcnt[i] : 0;
    s += v[r];
  }
  i32 x = s;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const i32 y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == WAVE - 1) sw[wv] = x;
  __syncthreads();
  i32 before = 0, all = 0;
  for (int w2 = 0; w2 < 1024 / WAVE; w2++) {
    if (w2 < wv) before += sw[w2];
    all += sw[w2];
  }
  i32 run = before + x - s;
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const int i = threadIdx.x * PER + r;
    if (i < nb) boff[i] = run;
    run += v[r];
  }
  if (threadIdx.x == 0) *total = all;
}
// Each block walks its range in tiles of CMP_TILE vertices: thread t of the
// block takes vertices t, t + CMP_BS, ... of the tile (coalesced), all loads
// of the tile are issued at once, and one scan of the 64 per-(round, wave)
// winner counts gives every winner its slot.
constexpr int CMP_PT = 16;
constexpr int CMP_TILE = CMP_BS * CMP_PT;
static_assert(CMP_PT * (CMP_BS / WAVE) == WAVE, "one scan lane per (round, wave)");
__global__ __launch_bounds__(CMP_BS) void k_scatter_moves(const u8* flag, i64 n, i64 per,
                                                         const i32* boff, const i32* part,
                                                         const i32* target, const float* gain,
                                                         i32 k, i32* mv_v, i32* mv_src,
                                                         i32* mv_dst, float* mv_gain,
                                                         i32* outc, i32* inc) {
  // per-part move counts of the block (k <= MAXK_LDS): out of src, into dst
  extern __shared__ i32 hsh[];
  __shared__ i32 scnt[WAVE];
  __shared__ i32 stot;
  constexpr int NW = CMP_BS / WAVE;
  const bool lds_h = k <= MAXK_LDS;
  i32* h = hsh;
  i32* h2 = hsh + (lds_h ? k : 0);
  if (lds_h)
    for (int q = threadIdx.x; q < k; q += CMP_BS) {
      h[q] = 0;
      h2[q] = 0;
    }
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const i64 b0 = (i64)blockIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  i64 run = boff[blockIdx.x];
  for (i64 sub = b0; sub < b1; sub += CMP_TILE) {
    bool f[CMP_PT];
    u64 bm[CMP_PT];
#pragma unroll
    for (int j = 0; j < CMP_PT; j++) {
      const i64 v = sub + (i64)j * CMP_BS + threadIdx.x;
      f[j] = v < b1 && flag[v];
    }
#pragma unroll
    for (int j = 0; j < CMP_PT; j++) bm[j] = __ballot(f[j]);
    __syncthreads();  // the previous tile's scan is read
    if (lane < CMP_PT) {
      i32 c = 0;
#pragma unroll
      for (int j = 0; j < CMP_PT; j++) c = lane == j ? __popcll(bm[j]) : c;
      scnt[lane * NW + wv] = c;
    }
    __syncthreads();
    if (wv == 0) {  // exclusive scan over (round, wave) in vertex order
      const i32 c = scnt[lane];
      i32 x = c;
#pragma unroll
      for (int o = 1; o < WAVE; o <<= 1) {
        const i32 y = __shfl_up(x, o);
        if (lane >= o) x += y;
      }
      scnt[lane] = x - c;
      if (lane == WAVE - 1) stot = x;
    }
    __syncthreads();
    i32 pp[CMP_PT], tt[CMP_PT];
    float gg[CMP_PT];
#pragma unroll
    for (int j = 0; j < CMP_PT; j++) {
      const i64 v = sub + (i64)j * CMP_BS + threadIdx.x;
      pp[j] = f[j] ? part[v] : 0;
      tt[j] = f[j] ? target[v] : 0;
      gg[j] = f[j] ? gain[v] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < CMP_PT; j++) {
      if (!f[j]) continue;
      const i64 v = sub + (i64)j * CMP_BS + threadIdx.x;
      const i64 p = run + scnt[j * NW + wv] + __popcll(bm[j] & lanes_below(lane));
      const i32 src = pp[j], dst = tt[j];
      mv_v[p] = (i32)v;
      mv_src[p] = src;
      mv_dst[p] = dst;
      mv_gain[p] = gg[j];
      if (lds_h) {
        atomicAdd(&h[src], 1);
        if (inc) atomicAdd(&h2[dst], 1);
      } else {
        atomicAdd(&outc[src], 1);
        if (inc) atomicAdd(&inc[dst], 1);
      }
    }
    run += stot;
  }
  __syncthreads();
  if (lds_h)
    for (int q = threadIdx.x; q < k; q += CMP_BS) {
      if (h[q]) atomicAdd(&outc[q], h[q]);
      if (inc && h2[q]) atomicAdd(&inc[q], h2[q]);
    }
}

// gpu_part.cu:357-369.
// With `inc` (moves per destination), *fit (set to 1 by k_bg_table) becomes 0
// when some destination has more moves than capacity: while it stays 1 every
// buffered move is selected whatever their order.
__global__ void k_capacity(const i32* pc, const i32* outc, i32 k, i64 n, int phase,
                           i64* avail, i32* zero1, const i32* inc, i32* fit) {
  if (zero1 && gtid() == 0) *zero1 = 0;
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
    if (inc && (i64)inc[q] > a) *fit = 0;
  }
}

// Selection when every move fits its destination (the capacity kernel's
// flag) and the network would have sorted all M moves ahead of the pads:
// all are selected, in buffer order (commit order does not change the result).
__global__ void k_select_all(i64 M, i32* sk_idx, u8* sel, i32* S, i32* pads) {
  for (i64 i = gtid(); i < M; i += gstride()) {
    sk_idx[i] = (i32)i;
    sel[i] = 1;
  }
  if (gtid() == 0) {
    *S = (i32)M;
    *pads = 0;
  }
}

// ------------------------------------------------------------ bitonic network --
// gpu_part.cu:472-506: the exact compare-exchange sequence on (gain, slot)
// pairs.  The element at the higher index of each pair compares itself with
// the lower one: swap = (gain[hi] < gain[lo]) ^ ((hi & kk) == 0); pairs with
// hi >= len are skipped.  Strides below the tile size (4096) run inside
// tiles held in registers (k_bt_tile); larger strides run as global passes
// that apply up to four consecutive strides at once (k_bt_global).
constexpr int BT_TB = 12;
constexpr int BT_SMAX = 5;

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

// value of the neighbouring lane (lane ^ 1), through DPP quad_perm [1,0,3,2]
__device__ __forceinline__ i32 lane_xor1_i(i32 v) {
  return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ float lane_xor1_f(float v) {
  return __int_as_float(lane_xor1_i(__float_as_int(v)));
}

// In-tile stages of the same network with 16 elements per thread held in
// registers under three layouts.  Layout R: register r of thread t holds tile
// position (t & (2^R - 1)) | r << R | (t >> R) << (R + 4), so the strides
// 2^R .. 2^(R+3) are exchanges between a thread's own registers.  A tile of
// 2^TB positions uses R = TB - 4 (A), R = TB - 8 (B) and R = 0 (C); with
// TB = 13 the one stride left (16) is an exchange between neighbouring lanes.
// Layouts change through LDS at swizzled addresses (p ^ ((p >> 4) & 63)),
// which keeps every layout's 32-bit accesses free of bank conflicts.
// Everything about a merge is a template argument: KB = log2 of the merge
// width when it is narrower than the tile (the direction then follows bit KB
// of the position), KB = 0 for a merge at least as wide as the tile (one
// direction for the whole tile, from its base); FULL = the tile lies wholly
// below len, so no pair is skipped.
template <int R>
__device__ __forceinline__ int bt_pos(int t, int r) {
  return (t & ((1 << R) - 1)) | (r << R) | ((t >> R) << (R + 4));
}
__device__ __forceinline__ int bt_swz(int p) { return p ^ ((p >> 4) & 63); }

template <int R, int JB, int KB, bool FULL, int UPC>
__device__ __forceinline__ void bt_tile_stage(float (&e)[16], i32 (&x)[16], int t, int lim,
                                              bool tile_up) {
  constexpr int J = 1 << JB;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    if (r & J) continue;
    const int rh = r | J;
    const int hi = bt_pos<R>(t, rh);
    const bool up = UPC >= 0 ? UPC != 0 : KB == 0 ? tile_up : ((hi >> KB) & 1) == 0;
    const float a = e[r], b = e[rh];
    const i32 ia = x[r], ib = x[rh];
    // branch-free exchange (selects, not a divergent branch)
    const bool sw = (FULL || hi < lim) && ((b < a) != up);
    e[r] = sw ? b : a;
    e[rh] = sw ? a : b;
    x[r] = sw ? ib : ia;
    x[rh] = sw ? ia : ib;
  }
}

// strides 2^BTOP down to 2^BBOT, all within R .. R + 3, in layout R
template <int R, int BTOP, int BBOT, int KB, bool FULL, int UPC>
__device__ __forceinline__ void bt_tile_stages(float (&e)[16], i32 (&x)[16], int t, int lim,
                                               bool tile_up) {
  if constexpr (BTOP >= BBOT) {
    bt_tile_stage<R, BTOP - R, KB, FULL, UPC>(e, x, t, lim, tile_up);
    bt_tile_stages<R, BTOP - 1, BBOT, KB, FULL, UPC>(e, x, t, lim, tile_up);
  }
}

template <int R, int RN>
__device__ __forceinline__ void bt_tile_relayout(float (&e)[16], i32 (&x)[16], float* sk, i32* si,
                                                 int t) {
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int q = bt_swz(bt_pos<R>(t, r));
    sk[q] = e[r];
    si[q] = x[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int q = bt_swz(bt_pos<RN>(t, r));
    e[r] = sk[q];
    x[r] = si[q];
  }
}

// One merge over the tile, entered in layout CUR (0: A, 2: C), left in C.
// UPC = 0 or 1: the direction of a merge wider than the tile, as a constant.
template <int TB, int KB, int CUR, bool FULL, int UPC = -1>
__device__ __forceinline__ void bt_tile_merge(float (&e)[16], i32 (&x)[16], float* sk, i32* si,
                                              int t, int lim, bool tile_up) {
  constexpr int RA = TB - 4, RB = TB - 8;
  constexpr int TOP = KB == 0 ? TB - 1 : KB - 1;  // highest stride bit of the merge
  if constexpr (TOP >= RA) {
    if constexpr (CUR == 2) bt_tile_relayout<0, RA>(e, x, sk, si, t);
    bt_tile_stages<RA, TOP, RA, KB, FULL, UPC>(e, x, t, lim, tile_up);
  }
  if constexpr (TOP >= RB) {
    if constexpr (TOP >= RA || CUR == 0) {
      bt_tile_relayout<RA, RB>(e, x, sk, si, t);
    } else {
      bt_tile_relayout<0, RB>(e, x, sk, si, t);
    }
    constexpr int BT = TOP < RA - 1 ? TOP : RA - 1;
    bt_tile_stages<RB, BT, RB, KB, FULL, UPC>(e, x, t, lim, tile_up);
    bt_tile_relayout<RB, 0>(e, x, sk, si, t);
  } else if constexpr (CUR == 0) {
    bt_tile_relayout<RA, 0>(e, x, sk, si, t);
  }
  if constexpr (TB == 13 && TOP >= 4) {  // stride 16: lanes t and t ^ 1
    const bool lo = (t & 1) == 0;
#pragma unroll
    for (int r = 0; r < 16; r++) {
      const float o = lane_xor1_f(e[r]);
      const i32 oi = lane_xor1_i(x[r]);
      const int p = bt_pos<0>(t, r);
      const int hi = lo ? p + 16 : p;
      const bool up = UPC >= 0 ? UPC != 0 : KB == 0 ? tile_up : ((hi >> KB) & 1) == 0;
      const float a = lo ? o : e[r];
      const float bb = lo ? e[r] : o;
      const bool sw = (FULL || hi < lim) && ((a < bb) != up);
      e[r] = sw ? o : e[r];
      x[r] = sw ? oi : x[r];
    }
  }
  bt_tile_stages<0, (TOP < 3 ? TOP : 3), 0, KB, FULL, UPC>(e, x, t, lim, tile_up);
}

// merges of width 2^KB, 2^(KB+1), ... up to the tile (and not above len)
template <int TB, int KB, bool FULL>
__device__ __forceinline__ void bt_tile_sort(float (&e)[16], i32 (&x)[16], float* sk, i32* si,
                                             int t, int lim, i64 base, i64 len) {
  if constexpr (KB <= TB) {
    if (((i64)1 << KB) > len) return;
    const bool tile_up = (base & ((i64)1 << KB)) == 0;
    bt_tile_merge<TB, (KB == TB ? 0 : KB), 2, FULL>(e, x, sk, si, t, lim, tile_up);
    bt_tile_sort<TB, KB + 1, FULL>(e, x, sk, si, t, lim, base, len);
  }
}

// FIRST: every merge up to the tile width (and not above len); otherwise the
// strides below the tile of the merge of width kk (wider than the tile).
// A partial last tile takes the FULL = false form, so the full tiles' code
// carries no bounds tests.
template <int TB, bool FULL, bool FIRST>
__device__ __forceinline__ void bt_tile_body(float* key, i32* idx, i64 len, i64 kk, float* sk,
                                             i32* si, const float* src, i64 M) {
  constexpr int TILE = 1 << TB;
  constexpr int RA = TB - 4;
  const i64 base = (i64)blockIdx.x * TILE;
  const int lim = FULL ? TILE : (int)(len - base);
  const int t = threadIdx.x;
  const float* kb = key + base;
  const i32* ib = idx + base;
  float e[16];
  i32 x[16];
  constexpr int RL = FIRST ? 0 : RA;  // load layout: C for a whole sort, A for a merge
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int p = bt_pos<RL>(t, r);
    if constexpr (FIRST) {  // the sort's input: moves [0, M), then (-1e9, -1) pads
      const i64 gp = base + p;
      e[r] = (FULL || p < lim) ? (gp < M ? src[gp] : -1e9f) : 0.f;
      x[r] = (FULL || p < lim) && gp < M ? (i32)gp : -1;
    } else {
      e[r] = (FULL || p < lim) ? kb[p] : 0.f;
      x[r] = (FULL || p < lim) ? ib[p] : -1;
    }
  }
  if constexpr (FIRST) bt_tile_sort<TB, 1, FULL>(e, x, sk, si, t, lim, base, len);
  else if ((base & kk) == 0) bt_tile_merge<TB, 0, 0, FULL, 1>(e, x, sk, si, t, lim, true);
  else bt_tile_merge<TB, 0, 0, FULL, 0>(e, x, sk, si, t, lim, false);
  const int o = bt_pos<0>(t, 0);
  float* ko = key + base;
  i32* io = idx + base;
  if constexpr (FULL) {
#pragma unroll
    for (int r = 0; r < 16; r += 4) {
      *reinterpret_cast<float4*>(ko + o + r) = make_float4(e[r], e[r + 1], e[r + 2], e[r + 3]);
      *reinterpret_cast<int4*>(io + o + r) = make_int4(x[r], x[r + 1], x[r + 2], x[r + 3]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < 16; r++) {
      if (o + r < lim) {
        ko[o + r] = e[r];
        io[o + r] = x[r];
      }
    }
  }
}

template <int TB, bool FIRST>
__global__ __launch_bounds__(1 << (TB - 4)) void k_bt_tile(float* key, i32* idx, i64 len,
                                                           i64 kk, const float* src, i64 M) {
  __shared__ float sk[1 << TB];
  __shared__ i32 si[1 << TB];
  if ((i64)(blockIdx.x + 1) << TB <= len)
    bt_tile_body<TB, true, FIRST>(key, idx, len, kk, sk, si, src, M);
  else
    bt_tile_body<TB, false, FIRST>(key, idx, len, kk, sk, si, src, M);
}

// k_bt_tile<TB, true> with M from the device: len = min(2^ceil(log2 M), n),
// the tiles below M rounded up to the tile (all: every tile below len).
template <int TB>
__global__ __launch_bounds__(1 << (TB - 4)) void k_bt_tile_first_dev(float* key, i32* idx,
                                                                     const i64* Mp, i64 n,
                                                                     const float* src, int all) {
  __shared__ float sk[1 << TB];
  __shared__ i32 si[1 << TB];
  constexpr i64 TILE = (i64)1 << TB;
  const i64 M = *Mp < n ? *Mp : n;
  if (M < 2) return;
  i64 p2 = 1;
  while (p2 < M) p2 <<= 1;
  const i64 len = p2 < n ? p2 : n;
  const i64 up = all ? len : ((M + TILE - 1) / TILE * TILE < len ? (M + TILE - 1) / TILE * TILE
                                                                 : len);
  if ((i64)blockIdx.x >= (up + TILE - 1) / TILE) return;
  if ((i64)(blockIdx.x + 1) << TB <= len)
    bt_tile_body<TB, true, true>(key, idx, len, 0, sk, si, src, M);
  else
    bt_tile_body<TB, false, true>(key, idx, len, 0, sk, si, src, M);
}

// S consecutive strides j, j/2, ..., j >> (S-1) of one merge kk in a single
// pass: every thread owns the 2^S elements whose positions differ only in
// those stride bits (a set closed under the S steps) and applies the steps in
// order in registers, which equals applying each step to the whole array.
template <int S>
__global__ __launch_bounds__(256) void k_bt_global(float* key, i32* idx, i64 len, i64 nsets,
                                                   i64 kk, int blo) {
  constexpr int R = 1 << S;
  const i64 lowmask = ((i64)1 << blo) - 1;
  for (i64 t = gtid(); t < nsets; t += gstride()) {
    const i64 base = ((t >> blo) << (blo + S)) | (t & lowmask);
    if (base >= len) continue;
    float a[R];
    i32 x[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const i64 p = base + ((i64)r << blo);
      if (p < len) {
        a[r] = key[p];
        x[r] = idx[p];
      }
    }
    const bool up = (base & kk) == 0;  // kk is above every stride bit
#pragma unroll
    for (int b = S - 1; b >= 0; b--) {
#pragma unroll
      for (int r = 0; r < R; r++) {
        if (r & (1 << b)) continue;
        const int rh = r | (1 << b);
        if (base + ((i64)rh << blo) < len) {
          if ((a[rh] < a[r]) ^ up) {
            const float tf = a[rh];
            a[rh] = a[r];
            a[r] = tf;
            const i32 ti = x[rh];
            x[rh] = x[r];
            x[r] = ti;
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      const i64 p = base + ((i64)r << blo);
      if (p < len) {
        key[p] = a[r];
        idx[p] = x[r];
      }
    }
  }
}


// ----------------------------------------------------------------- selection --
// gpu_part.cu:373-380 under SEQ: move at sorted position i is selected iff its
// rank among earlier positions with the same destination is below
// avail[dst].  Padding that reached the prefix is not selected.
constexpr int SEL_BLK = 1024;

// sdst[i] = destination of the i-th move in sorted order (-1 for a pad), and
// per-block destination counts H[k][nb].
__global__ __launch_bounds__(SEL_BLK) void k_sel_hist(const i32* sk_idx, const i32* mv_dst, i64 M,
                                                      i32 k, i32* H, i32* sdst) {
  __shared__ i32 h[MAXK_LDS];
  for (int q = threadIdx.x; q < k; q += SEL_BLK) h[q] = 0;
  __syncthreads();
  const i64 i = (i64)blockIdx.x * SEL_BLK + threadIdx.x;
  if (i < M) {
    const i32 id = sk_idx[i];
    const i32 d = id >= 0 ? mv_dst[id] : -1;
    sdst[i] = d;
    if (d >= 0 && d < k) atomicAdd(&h[d], 1);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < k; q += SEL_BLK) H[(i64)q * gridDim.x + blockIdx.x] = h[q];
}

// H is [k][nb]: one block per destination turns its row of per-block counts
// into exclusive prefix sums (the rank of each block's first move to q).
constexpr int SELSCAN_BS = 1024;
__global__ __launch_bounds__(SELSCAN_BS) void k_sel_scan(i32* H, i64 nb) {
  __shared__ i32 s_w[SELSCAN_BS / WAVE];
  i32* row = H + (i64)blockIdx.x * nb;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  i32 carry = 0;
  for (i64 base = 0; base < nb; base += SELSCAN_BS) {
    const i64 b = base + threadIdx.x;
    const i32 v = b < nb ? row[b] : 0;
    i32 x = v;  // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const i32 y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) s_w[wv] = x;
    __syncthreads();
    i32 before = carry, total = carry;
    for (int w2 = 0; w2 < SELSCAN_BS / WAVE; w2++) {
      if (w2 < wv) before += s_w[w2];
      total += s_w[w2];
    }
    if (b < nb) row[b] = before + x - v;
    carry = total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(SEL_BLK) void k_sel_mark(const i32* sdst, i64 M, i32 k,
                                                      const i32* H, const i64* avail, u8* sel,
                                                      i32* S, i32* pads) {
  __shared__ i32 cntd[MAXK_LDS];
  __shared__ i32 ssel[SEL_BLK / WAVE];
  for (int q = threadIdx.x; q < k; q += SEL_BLK) cntd[q] = H[(i64)q * gridDim.x + blockIdx.x];
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

// Ranks within the block in position order: every wave ranks its own moves
// per destination (one ballot round per distinct destination in the wave) and
// records its per-destination counts; a scan over the block's waves per
// destination then gives each wave's first rank (k <= SELW_MAXK).
constexpr int SELW_MAXK = 256;
__global__ __launch_bounds__(SEL_BLK) void k_sel_mark_w(const i32* sdst, i64 M, i32 k,
                                                      const i32* H, const i64* avail, u8* sel,
                                                      i32* S, i32* pads) {
  constexpr int NW = SEL_BLK / WAVE;
  __shared__ i32 cntd[SELW_MAXK];
  __shared__ i32 wcnt[NW][SELW_MAXK];  // per wave and destination; then the wave's first rank
  __shared__ i32 ssel[NW];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int q = threadIdx.x; q < k; q += SEL_BLK) cntd[q] = H[(i64)q * gridDim.x + blockIdx.x];
  for (int q = lane; q < k; q += WAVE) wcnt[wv][q] = 0;
  const i64 i = (i64)blockIdx.x * SEL_BLK + threadIdx.x;
  const i32 d = i < M ? sdst[i] : -1;
  const bool valid = d >= 0 && d < k;
  i32 inwave = 0;  // rank among this wave's earlier moves to d
  __builtin_amdgcn_wave_barrier();
  u64 active = __ballot(valid);
  while (active) {
    const int leader = __ffsll((unsigned long long)active) - 1;
    const i32 dl = __shfl(d, leader);
    const u64 same = __ballot(valid && d == dl);
    if (valid && d == dl) inwave = __popcll(same & lanes_below(lane));
    if (lane == leader) wcnt[wv][dl] = __popcll(same);
    active &= ~same;
  }
  __syncthreads();
  // exclusive scan over the waves, per destination, from the block's base
  for (int q = threadIdx.x; q < k; q += SEL_BLK) {
    i32 run = cntd[q];
#pragma unroll
    for (int w = 0; w < NW; w++) {
      const i32 c = wcnt[w][q];
      wcnt[w][q] = run;
      run += c;
    }
  }
  __syncthreads();
  bool mine = false;
  if (valid) mine = (i64)wcnt[wv][d] + inwave < avail[d];
  if (i < M) sel[i] = mine ? 1 : 0;
  const u64 pm = __ballot(i < M && !valid);
  if (lane == 0 && pm) atomicAdd(pads, (i32)__popcll(pm));
  const u64 bm = __ballot(mine);
  if (lane == 0) ssel[wv] = __popcll(bm);
  __syncthreads();
  if (threadIdx.x == 0) {
    i32 s = 0;
    for (int w = 0; w < NW; w++) s += ssel[w];
    if (s) atomicAdd(S, s);
  }
}

// part8[v] = part[v] when it names a part, 255 otherwise (k <= PART8_MAXK)
__global__ void k_part8(const i32* part, i64 n, i32 k, u8* part8) {
  for (i64 i = gtid(); i < n; i += gstride()) {
    const i32 p = part[i];
    part8[i] = (p >= 0 && p < k) ? (u8)p : (u8)255;
  }
}

// gpu_part.cu:383-397.  Every vertex has at most one move, so the commits are
// independent; the count changes are summed per block in LDS (k <= MAXK_LDS)
// and added to pc once per block and part.
template <bool LDS>
__global__ __launch_bounds__(256) void k_commit(const i32* sk_idx, const u8* sel, i64 M,
                                                const i32* mv_v, const i32* mv_src,
                                                const i32* mv_dst, i32* part, u8* part8,
                                                i32* pc, i32 k, i32* app, i32* napp) {
  __shared__ i32 dpc[LDS ? MAXK_LDS : 1];
  if (LDS) {
    for (int q = threadIdx.x; q < k; q += blockDim.x) dpc[q] = 0;
    __syncthreads();
  }
  const int lane = lane_id();
  for (i64 base = (i64)blockIdx.x * blockDim.x; base < M; base += gstride()) {
    const i64 i = base + threadIdx.x;
    bool applied = false;
    i32 id = -1;
    if (i < M && sel[i]) {
      id = sk_idx[i];
      if (id >= 0) {
        const i32 src = mv_src[id], dst = mv_dst[id];
        if (src < k && dst < k && src >= 0 && dst >= 0) {
          const i32 v = mv_v[id];
          const i32 old = part[v];
          if (old == src) {
            part[v] = dst;
            if (part8) part8[v] = (u8)dst;
            atomicAdd(LDS ? &dpc[dst] : &pc[dst], 1);
            atomicAdd(LDS ? &dpc[src] : &pc[src], -1);
            applied = true;
          }
        }
      }
    }
    if (app) {  // the applied moves, for the tally update (any order)
      const u64 am = __ballot(applied);
      i32 ob = 0;
      if (lane == 0 && am) ob = atomicAdd(napp, (i32)__popcll(am));
      ob = __shfl(ob, 0);
      if (applied) app[ob + __popcll(am & lanes_below(lane))] = id;
    }
  }
  if (LDS) {
    __syncthreads();
    for (int q = threadIdx.x; q < k; q += blockDim.x)
      if (dpc[q]) atomicAdd(&pc[q], dpc[q]);
  }
}

// -------------------------------------------------------------------- cut --
// gpu_part.cu:400-422: sum of w over arcs whose ends lie in different parts.
// Summed in uint64 (the int32 wrap of the reference is its low 32 bits).

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
// PT = u8: the byte copy part8 (k <= 255, 255 for every value outside
// [0, k)), whose gathers stay in L2; the test is the same.
template <bool UNIT, typename PT>
__global__ __launch_bounds__(256) void k_cut_arcs(const i32* asrc, const i32* adj, const i32* w,
                                                  i64 n, i64 e_end, const PT* part, i32 k,
                                                  u64* acc) {
  u64 s = 0;
  for (i64 i = gtid(); i < e_end; i += gstride()) {
    const i32 v = asrc[i];
    const i32 u = adj[i];
    if (u < n) {
      const i32 pv = part[v], pu = part[u];
      const bool in = sizeof(PT) == 1 ? (pv != 255 && pu != 255)
                                      : (pv >= 0 && pv < k && pu >= 0 && pu < k);
      if (pv != pu && in) s += UNIT ? 1ull : (u64)(i64)w[i];
    }
  }
  block_add_u64(s, acc);
}

// One word block with a pass's host-side inputs, copied to the host at once:
// out[0] = *m, out[1] = *cut, out32[4..5] = ctr[0..1], out32[6..6+k) = pc
// (absent inputs leave zeros).
// out: [M (i32), fit (i32)], cut, S, pads, pc[k]
__global__ __launch_bounds__(256) void k_pass_status(const i64* m, const u64* cut,
                                                     const i32* ctr, const i32* pc, i32 k,
                                                     const i32* fit, i64* out) {
  i32* o32 = (i32*)out;
  if (threadIdx.x == 0) {
    o32[0] = m ? (i32)*m : 0;
    o32[1] = fit ? *fit : 0;
    out[1] = cut ? (i64)*cut : 0;
    o32[4] = ctr ? ctr[0] : 0;
    o32[5] = ctr ? ctr[1] : 0;
  }
  if (pc)
    for (int q = threadIdx.x; q < k; q += blockDim.x) o32[6 + q] = pc[q];
}

// *out = sum of the accumulator slots; the slots are cleared for the next use.
__global__ __launch_bounds__(ACC_SLOTS) void k_acc_finish(u64* acc, u64* out) {
  u64* slot = &acc[threadIdx.x * ACC_STRIDE];
  u64 v = *slot;
  *slot = 0;
  __shared__ u64 sw[ACC_SLOTS / WAVE];
  v = wave_sum(v);
  if (lane_id() == 0) sw[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < ACC_SLOTS / WAVE; w++) t += sw[w];
    *out = t;
  }
}

#include "tgp_batch_engine.inc"

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
                         i32* heavy_count, i32* nonunit, u64* maxabs, hipStream_t s) {
  k_level_census<<<grid_for(g.n > g.m ? g.n : g.m, 256, 8192), 256, 0, s>>>(
      g, nedges, heavy_deg, heavy_list, heavy_count, nonunit, maxabs);
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
  k_fill<i64><<<1, 1, 0, s>>>(lo + 3, 1, n);  // first-undecided accumulator
}

void launch_chain_lists(DGraph g, i64 minlen, i32* lidx, i32* lrows, i32* nl, i32* lists,
                        i32* lcnt, hipStream_t s) {
  k_list_rows<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, minlen, lidx, lrows, nl);
  k_long_lists<<<grid_for(g.m / (minlen > 0 ? minlen : 1) + 1, 1, 1u << 14), LST_BS, 0, s>>>(
      g, lrows, nl, lists, lcnt);
}

void launch_match_chain(DGraph g, ChainLow low, i32* st, i32* blkA, i32* blkC, i32* prop,
                        const i64* lo, u32* ticket, i64 budget, int blocks, int sg, int per,
                        u64* stats, hipStream_t s) {
  // sub-group width (lanes per vertex) and register entries per lane: rows
  // of up to sg * per arcs are held in registers
  // blocks <= 0: as many as the device holds at once (every wave then takes
  // part in the chain from the start; more blocks only queue behind them)
  auto resident = [&](const void* fn) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      (void)hipGetLastError();
      return 512;
    }
    return std::max(1, per_cu * cus);
  };
  // the event counts (stats) are kept by the two default shapes' counting
  // forms only
#define TGP_CHAIN_CASE(SG, PER, ST)                                                           \
  if (sg == SG && per == PER && (ST == (stats != nullptr) || !ST)) {                          \
    const int nb = blocks > 0 ? blocks : resident((const void*)k_match_chain<SG, PER, ST>);   \
    k_match_chain<SG, PER, ST><<<nb, 256, 0, s>>>(g, low, st, blkA, blkC, prop, lo, ticket,   \
                                                  budget, stats);                             \
    return;                                                                                   \
  }
  TGP_CHAIN_CASE(16, 8, true)
  TGP_CHAIN_CASE(8, 8, true)
  TGP_CHAIN_CASE(16, 8, false)
  TGP_CHAIN_CASE(32, 8, false)
#undef TGP_CHAIN_CASE
  const int nb = blocks > 0 ? blocks : resident((const void*)k_match_chain<8, 8, false>);
  k_match_chain<8, 8, false><<<nb, 256, 0, s>>>(g, low, st, blkA, blkC, prop, lo, ticket, budget,
                                                stats);
}

void launch_chain_low_count(DGraph g, i64 minlen, i32* list, i32* nlist, i64* cnt,
                            hipStream_t s) {
  k_long_rows<<<grid_for(g.n, 256, 1u << 14), 256, 0, s>>>(g, minlen, list, nlist);
  k_chain_low<false><<<8192, 256, 0, s>>>(g, list, nlist, cnt, nullptr, nullptr);
}
void launch_chain_low_write(DGraph g, const i32* list, const i32* nlist, const i64* off, i32* out,
                            hipStream_t s) {
  k_chain_low<true><<<8192, 256, 0, s>>>(g, list, nlist, nullptr, off, out);
}

void launch_max_row(DGraph g, u64* out, hipStream_t s) {
  k_max_row<<<grid_for(g.n, 256, 4096), 256, 0, s>>>(g, out);
}
void launch_first_undecided(const i32* st, i64 n, i64* lo, hipStream_t s) {
  k_first_undecided_st<<<grid_for(n, 256 * 16, 4096), 256, 0, s>>>(st, n, lo);
  k_first_undecided_take<<<1, 1, 0, s>>>(lo, n);
}

void launch_chain_export(const i32* st, i64 n, i32* mate, u8* dec, hipStream_t s) {
  k_chain_export<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(st, n, mate, dec);
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

void launch_rows_strict(DGraph g, i64 m, i32* crow, i32* flags, hipStream_t s) {
  const i64 nch = sym_check_chunks(m);
  if (nch <= 0) return;
  k_chunk_rows<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, nch, crow);
  k_rows_strict<<<grid_for(nch, 4, 1u << 16), 256, 0, s>>>(g, crow, nch, m, flags);
}
void launch_csr_check(const i64* xadj, const i32* adj, i64 n, i64 m, i32* flag, hipStream_t s) {
  k_csr_check<<<grid_for(std::max(n + 1, m), 256, 8192), 256, 0, s>>>(xadj, adj, n, m, flag);
}

void launch_batch_csr_check(const BatchGraph* dg, const i32* list, i32 cnt, i32* bad,
                            hipStream_t s) {
  if (cnt > 0) k_batch_csr_check<<<grid_for(cnt, 1, 4096), 256, 0, s>>>(dg, list, cnt, bad);
}

i64 sym_check_chunks(i64 m) { return (m + SYM_CHUNK - 1) / SYM_CHUNK; }
void launch_sym_check_sorted(DGraph g, i64 m, u8* cls, const i32* crow, i32* flags, u64* cnt,
                             hipStream_t s) {
  k_deg_class<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, cls);
  const i64 nch = sym_check_chunks(m);
  if (nch <= 0) return;
  k_sym_search_sorted<<<grid_for(nch, 4, 1u << 16), 256, 0, s>>>(g, cls, crow, nch, m, flags,
                                                                 cnt);
}
void launch_sym_check(DGraph g, u64* table, i32* wt, i64 cap, i32* flags, hipStream_t s) {
  k_fill<u64><<<grid_for(cap, 256, 1u << 16), 256, 0, s>>>(table, cap, ~0ull);
  k_sym_insert<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, wt, cap, flags);
  k_sym_check<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, wt, cap, flags);
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

void launch_rep_classify(const i64* xadj, const i32* tb, i64 n_c, i64* dsz, i64* dlong,
                         i64* csz, i32* med, i32* nmed, i32* lng, i32* nlng, i64* nchk,
                         i32* meds, i32* nmeds, hipStream_t s) {
  k_rep_classify<<<grid_for(n_c, 256, 1u << 16), 256, 0, s>>>(
      xadj, tb, n_c, dsz, dlong, csz, med, nmed, lng, nlng, nchk, meds, nmeds);
}
void launch_rep_chunks(const i32* lng, i64 nl, const i64* dsz, const i64* cbase, i32* ch_row,
                       i64* ch_beg, hipStream_t s) {
  if (nl > 0)
    k_rep_chunks<<<grid_for(nl, 256, 4096), 256, 0, s>>>(lng, nl, dsz, cbase, ch_row, ch_beg);
}
void launch_rep_long(DGraph g, const i32* tb, const i32* new_ids, const RepLong& L, int pass,
                     const i64* ox, i32* oa, i32* ow, hipStream_t s) {
  if (L.nch <= 0) return;
  const unsigned gr = grid_for(L.nch, 1, 1u << 16);
  if (pass == 0) {
    k_rep_insert<<<gr, 256, 0, s>>>(g, tb, new_ids, L.lng, L.ch_row, L.ch_beg, L.nch, L.dsz,
                                    L.toff, L.csz, L.T, L.W);
    k_rep_mark<<<gr, 256, 0, s>>>(g, tb, new_ids, L.lng, L.ch_row, L.ch_beg, L.nch, L.dsz,
                                  L.toff, L.csz, L.eoff, L.T, L.E, L.chunk_cnt);
  } else {
    k_rep_emit<<<gr, 256, 0, s>>>(L.lng, L.ch_row, L.ch_beg, L.nch, L.dsz, L.cbase, L.choff,
                                  L.toff, L.eoff, L.T, L.W, L.E, ox, oa, ow);
  }
}
void launch_rl_buckets(const RepLong& L, i64 target, i32* lgb, i64* nbk, hipStream_t s) {
  if (L.nl > 0)
    k_rl_buckets<<<grid_for(L.nl, 256, 4096), 256, 0, s>>>(L.lng, L.nl, L.dsz,
                                                           target > 0 ? target : RL_BK_TARGET,
                                                           lgb, nbk);
}
void launch_rl_brow(i64 nl, const i64* bbase, i32* brow, hipStream_t s) {
  if (nl > 0) k_rl_brow<<<grid_for(nl, 256, 4096), 256, 0, s>>>(nl, bbase, brow);
}
void launch_rl_fetch(DGraph g, const i32* tb, const i32* new_ids, const RepLong& L,
                     const RepBuckets& B, hipStream_t s) {
  if (L.nch <= 0) return;
  k_rl_fetch<<<grid_for(L.nch, 1, 1u << 16), 256, 0, s>>>(g, tb, new_ids, L.lng, L.ch_row,
                                                          L.ch_beg, L.nch, L.dsz, L.eoff, B.lgb,
                                                          B.bbase, B.R, B.bcnt);
}
void launch_rl_scatter(const RepLong& L, const RepBuckets& B, hipStream_t s) {
  if (L.nch <= 0) return;
  k_rl_scatter<<<grid_for(L.nch, 1, 1u << 16), 256, 0, s>>>(L.lng, L.ch_row, L.ch_beg, L.nch,
                                                            L.dsz, L.eoff, B.lgb, B.bbase,
                                                            B.boff, B.bcur, B.R, B.K, B.W);
}
void launch_rl_dedup(const RepLong& L, const RepBuckets& B, i32 fill, hipStream_t s) {
  if (B.nbt <= 0) return;
  fill = std::max<i32>(1, std::min<i32>(fill, RL_TFILL));
  k_rl_dedup<<<grid_for(B.nbt, 1, 1u << 16), 256, 0, s>>>(B.boff, B.nbt, B.brow, L.lng, L.eoff,
                                                          B.K, B.W, B.R, fill, B.over);
}
void launch_rl_count(const RepLong& L, const RepBuckets& B, hipStream_t s) {
  if (L.nch <= 0) return;
  k_rl_count<<<grid_for(L.nch, 1, 1u << 16), 256, 0, s>>>(L.lng, L.ch_row, L.ch_beg, L.nch,
                                                          L.dsz, L.eoff, B.R, L.chunk_cnt);
}
void launch_rl_emit(const RepLong& L, const RepBuckets& B, const i64* ox, i32* oa, i32* ow,
                    hipStream_t s) {
  if (L.nch <= 0) return;
  k_rl_emit<<<grid_for(L.nch, 1, 1u << 16), 256, 0, s>>>(L.lng, L.ch_row, L.ch_beg, L.nch, L.dsz,
                                                         L.cbase, L.choff, L.eoff, B.R, ox, oa,
                                                         ow);
}
void launch_rep_rowcnt(const RepLong& L, i64* cnt, hipStream_t s) {
  if (L.nl > 0)
    k_rep_rowcnt<<<grid_for(L.nl, 256, 4096), 256, 0, s>>>(L.lng, L.nl, L.cbase, L.choff, cnt);
}
void launch_rep_rows(DGraph g, const i32* tb, const i32* new_ids, i64 n_c, const i64* dsz,
                     const i32* med, i64 nmed, const i32* meds, i64 nmeds, const i64* eoff,
                     u64* R, int pass, i64* cnt, const i64* ox, i32* oa, i32* ow,
                     hipStream_t s) {
  const unsigned g8 = grid_for(n_c, 256 / 8, 1u << 16), g64 = grid_for(n_c, 4, 1u << 16);
  const unsigned gm = grid_for(nmed, 1, 1u << 16), gw = grid_for(nmeds, 4, 1u << 16);
  if (pass == 0) {
    k_rep_small<8, 0><<<g8, 256, 0, s>>>(g, tb, new_ids, n_c, dsz, 0, cnt, ox, oa, ow);
    k_rep_small<64, 0><<<g64, 256, 0, s>>>(g, tb, new_ids, n_c, dsz, 9, cnt, ox, oa, ow);
    if (nmeds > 0)
      k_rep_medw0<<<gw, 256, 0, s>>>(g, tb, new_ids, meds, nmeds, dsz, eoff, R, cnt);
    if (nmed > 0) k_rep_med0<<<gm, 256, 0, s>>>(g, tb, new_ids, med, nmed, dsz, eoff, R, cnt);
  } else {
    k_rep_small<8, 1><<<g8, 256, 0, s>>>(g, tb, new_ids, n_c, dsz, 0, cnt, ox, oa, ow);
    k_rep_small<64, 1><<<g64, 256, 0, s>>>(g, tb, new_ids, n_c, dsz, 9, cnt, ox, oa, ow);
    if (nmeds > 0) k_rep_medw1<<<gw, 256, 0, s>>>(meds, nmeds, dsz, eoff, R, ox, oa, ow);
    if (nmed > 0) k_rep_med1<<<gm, 256, 0, s>>>(med, nmed, dsz, eoff, R, ox, oa, ow);
  }
}
void launch_pairs(DGraph g, u64* table, i64 cap, u64* count, hipStream_t s) {
  k_pairs<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, table, cap, count);
}
void launch_count_pair_arcs(DGraph g, u64* count, hipStream_t s) {
  k_count_pair_arcs<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, count);
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
                     i32* zero_k, i32 nzero, i32* zero_2, i32* one1, hipStream_t s) {
  k_bg_table<<<grid_for(std::max<i64>(k, nzero), 256, 64), 256, 0, s>>>(
      pc, k, n, phase, bg, src_big, zero_k, nzero, zero_2, one1);
}

i64 gain_gscratch_elems(i32 k) { return k <= MAXK_LDS ? 0 : (i64)65536 * k; }

void launch_split_heavy(DGraph g, i64 nedges, const i32* heavy, i64 nheavy, i32* mid,
                        i32* nmid, i32* medium, i32* nmedium, i32* huge, i32* nhuge,
                        hipStream_t s) {
  if (nheavy > 0)
    k_split_heavy<<<grid_for(nheavy, 256, 4096), 256, 0, s>>>(g, nedges, heavy, nheavy, mid,
                                                              nmid, medium, nmedium, huge, nhuge);
}

void launch_huge_bounds(DGraph g, i64 nedges, const i32* huge, i64 nhuge, i64* hb,
                        hipStream_t s) {
  if (nhuge > 0)
    k_huge_bounds<<<grid_for(nhuge, 256, 1024), 256, 0, s>>>(g, nedges, huge, nhuge, hb);
}

void launch_row_abs(DGraph g, i64 nedges, i64 heavy_deg, const i32* heavy, i64 nheavy,
                    const u64* maxabs, i32* over, hipStream_t s) {
  k_row_abs_light<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, nedges, heavy_deg, maxabs,
                                                                over);
  if (nheavy > 0)
    k_row_abs_heavy<<<grid_for(nheavy, 4, 8192), 256, 0, s>>>(g, nedges, heavy, nheavy, maxabs,
                                                               over);
}

template <typename PT>
static void launch_gain_small_k(DGraph g, i64 nedges, const PT* part, const i32* part32, i32 k,
                                const float* bg, const u8* src_big, float* gain, i32* target,
                                float* cand, int phase, int unit, const GainLists& L,
                                u32* ctr, u64* cut_acc, TallyRef Tout, hipStream_t s) {
  const size_t shm = (size_t)(2 * k + 4 * 8 * (k | 1)) * 4;
  k_fill<u32><<<1, 256, 0, s>>>(ctr, GAIN_CTRS * 32, 0u);
  if (unit)
    k_gain_fused<true, PT><<<GAIN_FUSED_BLOCKS, 256, shm, s>>>(g, nedges, part, k, bg, src_big,
                                                               gain, target, cand, phase, L,
                                                               ctr, cut_acc, Tout);
  else
    k_gain_fused<false, PT><<<GAIN_FUSED_BLOCKS, 256, shm, s>>>(g, nedges, part, k, bg, src_big,
                                                                gain, target, cand, phase, L,
                                                                ctr, cut_acc, Tout);
  if (L.nhuge > 0) {
    const size_t shmf = (size_t)(3 * k) * 4;
    const unsigned grf = grid_for(L.nhuge, 1, 8192);
    if (unit)
      k_gain_huge_final<true><<<grf, 256, shmf, s>>>(g, nedges, part32, k, bg, src_big, gain,
                                                     target, cand, phase, L.huge, L.nhuge,
                                                     L.gtally, L.gabs, Tout);
    else
      k_gain_huge_final<false><<<grf, 256, shmf, s>>>(g, nedges, part32, k, bg, src_big, gain,
                                                      target, cand, phase, L.huge, L.nhuge,
                                                      L.gtally, L.gabs, Tout);
  }
}

bool launch_gain(DGraph g, i64 nedges, const i32* part, const u8* part8, i32 k, const float* bg,
                 const u8* src_big, float* gain, i32* target, float* cand, int phase, int unit,
                 i32* gscratch, const GainLists& L, u64* cut_acc, const TallyRef* Tout,
                 hipStream_t s) {
  const TallyRef tout = Tout ? *Tout : TallyRef{};
  if (k <= 512) {
    if (part8 && k <= PART8_MAXK)
      launch_gain_small_k<u8>(g, nedges, part8, part, k, bg, src_big, gain, target, cand, phase,
                              unit, L, L.ctr, cut_acc, tout, s);
    else
      launch_gain_small_k<i32>(g, nedges, part, part, k, bg, src_big, gain, target, cand, phase,
                               unit, L, L.ctr, cut_acc, tout, s);
    return cut_acc != nullptr;
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
  return false;  // the large-k kernels do not tally the cut
}

void launch_gain_rows(DGraph g, i64 nedges, const i32* part, i32 k, const float* bg,
                      const u8* src_big, float* gain, i32* target, float* cand, int phase,
                      int unit, const i32* list, i64 cnt, hipStream_t s) {
  if (cnt <= 0) return;
  const size_t shm = (size_t)k * 12;
  const unsigned gr = grid_for(cnt, 1, 65536);
  if (unit)
    k_gain<true, true><<<gr, WAVE, shm, s>>>(g, nedges, part, k, bg, src_big, gain, target, cand,
                                             phase, nullptr, list, cnt, nullptr, nullptr);
  else
    k_gain<false, true><<<gr, WAVE, shm, s>>>(g, nedges, part, k, bg, src_big, gain, target,
                                              cand, phase, nullptr, list, cnt, nullptr, nullptr);
}

void launch_token(DGraph g, i64 nedges, const float* cand, const i32* part, const i32* target,
                  i32 k, u8* flag, const i32* heavy_list, i64 heavy_count, i64 heavy_deg,
                  int lanes, i32* wit, hipStream_t s) {
  const int G = (lanes == 0 || lanes == 4 || lanes == 8) ? lanes : 1;
  const unsigned bl = grid_for(g.n, G ? 256 / G : 256, 1u << 16);
  const unsigned bh = heavy_count > 0 ? grid_for(heavy_count, 4, 8192) : 0;
  if (G == 0)
    k_token<0><<<bl + bh, 256, 0, s>>>(g, nedges, cand, part, target, k, flag, heavy_deg,
                                       heavy_list, heavy_count, bh, wit);
  else if (G == 8)
    k_token<8><<<bl + bh, 256, 0, s>>>(g, nedges, cand, part, target, k, flag, heavy_deg,
                                       heavy_list, heavy_count, bh, wit);
  else if (G == 4)
    k_token<4><<<bl + bh, 256, 0, s>>>(g, nedges, cand, part, target, k, flag, heavy_deg,
                                       heavy_list, heavy_count, bh, wit);
  else
    k_token<1><<<bl + bh, 256, 0, s>>>(g, nedges, cand, part, target, k, flag, heavy_deg,
                                       heavy_list, heavy_count, bh, wit);
}

void launch_compact_moves(const u8* flag, i64 n, const i32* part, const i32* target,
                          const float* gain, i32 k, i32* bcnt, i32* boff, i64* total,
                          i32* mv_v, i32* mv_src, i32* mv_dst, float* mv_gain, i32* outc,
                          i32* inc, hipStream_t s) {
  i64 nb = (n + 4095) / 4096;
  if (nb < 1) nb = 1;
  if (nb > CMP_MAXB) nb = CMP_MAXB;
  i64 per = (n + nb - 1) / nb;
  per = (per + CMP_TILE - 1) / CMP_TILE * CMP_TILE;
  nb = (n + per - 1) / per;
  if (nb < 1) nb = 1;
  k_flag_count<<<(unsigned)nb, CMP_BS, 0, s>>>(flag, n, per, bcnt);
  k_flag_offsets<<<1, 1024, 0, s>>>(bcnt, (i32)nb, boff, total);
  const size_t lds = k <= MAXK_LDS ? (size_t)2 * k * sizeof(i32) : 0;
  k_scatter_moves<<<(unsigned)nb, CMP_BS, lds, s>>>(flag, n, per, boff, part, target, gain, k,
                                                     mv_v, mv_src, mv_dst, mv_gain, outc, inc);
}

void launch_capacity(const i32* pc, const i32* outc, i32 k, i64 n, int phase, i64* avail,
                     i32* zero1, const i32* inc, i32* fit, hipStream_t s) {
  k_capacity<<<grid_for(k, 256, 64), 256, 0, s>>>(pc, outc, k, n, phase, avail, zero1, inc, fit);
}

void launch_select_all(i64 M, i32* sk_idx, u8* sel, i32* S, i32* pads, hipStream_t s) {
  k_select_all<<<grid_for(M, 256 * 4, 1u << 14), 256, 0, s>>>(M, sk_idx, sel, S, pads);
}

static bool sort_all_env() {
  const char* all_env = std::getenv("TGP_SORT_ALL");
  return all_env && std::atoi(all_env) != 0;
}

void launch_sort_first_dev(const float* mv_gain, const i64* Mp, i64 n, float* key, i32* idx,
                           hipStream_t s) {
  constexpr i64 TILE = (i64)1 << BT_TB;
  constexpr int NT = 1 << (BT_TB - 4);
  if (n < 2) return;
  k_bt_tile_first_dev<BT_TB><<<(unsigned)((n + TILE - 1) / TILE), NT, 0, s>>>(
      key, idx, Mp, n, mv_gain, sort_all_env() ? 1 : 0);
}

void launch_sort_moves(const float* mv_gain, i64 M, i64 len, float* key, i32* idx,
                       hipStream_t s, bool first_done) {
  constexpr i64 TILE = (i64)1 << BT_TB;
  if (len < 2) {
    k_sort_init<<<1, 64, 0, s>>>(mv_gain, M, len, key, idx);
    return;
  }
  const unsigned tiles = (unsigned)((len + TILE - 1) / TILE);
  constexpr int NT = 1 << (BT_TB - 4);
  // Positions [M, len) hold identical (-1e9, -1) pads, and within a merge
  // block every comparator acts on that block alone, so a block lying wholly
  // at or past M compares pads with pads and changes nothing: each merge of
  // size kk runs only on the blocks below M rounded up to kk (nothing is
  // skipped under TGP_SORT_ALL=1).  The first pass reads the moves itself and
  // writes the pads of its tiles; the tiles past them take the pads by a fill.
  const char* smax_env = std::getenv("TGP_SORT_SMAX");  // strides per global pass (1-5)
  const int smax = smax_env ? std::max(1, std::min(BT_SMAX, std::atoi(smax_env))) : BT_SMAX;
  const bool all = sort_all_env();
  auto upto = [&](i64 kk) { return all ? len : std::min(len, (M + kk - 1) / kk * kk); };
  const unsigned t0 = (unsigned)((upto(TILE) + TILE - 1) / TILE);
  if (!first_done) k_bt_tile<BT_TB, true><<<t0, NT, 0, s>>>(key, idx, len, 0, mv_gain, M);
  const i64 pad0 = (i64)t0 * TILE;
  if (pad0 < len) {
    const unsigned gf = grid_for(len - pad0, 256, 1u << 16);
    k_fill<float><<<gf, 256, 0, s>>>(key + pad0, len - pad0, -1e9f);
    k_fill<i32><<<gf, 256, 0, s>>>(idx + pad0, len - pad0, -1);
  }
  i64 p2 = 1;
  while (p2 < len) p2 <<= 1;
  for (i64 kk = 2 * TILE; kk <= len; kk <<= 1) {
    const i64 mk = upto(kk);
    int bhi = 0;  // log2 of the largest stride of this merge
    while (((i64)1 << (bhi + 1)) < kk) bhi++;
    for (int b = bhi; b >= BT_TB;) {
      // Up to 5 strides per pass in registers (32 elements per thread; 4:
      // 171 vs 165 ms, 6: 170 ms per partition on R-MAT 24).  Staging 5-8
      // strides in LDS measured slower than registers.
      const int S = std::min(smax, b - BT_TB + 1);
      const int blo = b - S + 1;
      // set t starts at ((t >> blo) << (blo + S)) | (t & (2^blo - 1)): the sets
      // below mk are the first ones
      const i64 span = (i64)1 << (blo + S);
      const i64 nsets = std::min(p2 >> S, ((mk + span - 1) / span) << blo);
      const unsigned gr = grid_for(nsets, 256, 1u << 16);
      if (S == 5) k_bt_global<5><<<gr, 256, 0, s>>>(key, idx, len, nsets, kk, blo);
      else if (S == 4) k_bt_global<4><<<gr, 256, 0, s>>>(key, idx, len, nsets, kk, blo);
      else if (S == 3) k_bt_global<3><<<gr, 256, 0, s>>>(key, idx, len, nsets, kk, blo);
      else if (S == 2) k_bt_global<2><<<gr, 256, 0, s>>>(key, idx, len, nsets, kk, blo);
      else k_bt_global<1><<<gr, 256, 0, s>>>(key, idx, len, nsets, kk, blo);
      b -= S;
    }
    const unsigned tk = (unsigned)((mk + TILE - 1) / TILE);
    k_bt_tile<BT_TB, false><<<tk, NT, 0, s>>>(key, idx, len, kk, nullptr, 0);
  }
}

i64 select_scratch_elems(i64 M, i32 k) { return ((M + SEL_BLK - 1) / SEL_BLK + 1) * (i64)k; }

void launch_select(const i32* sk_idx, const i32* mv_dst, i64 M, i32 k, const i64* avail,
                   i32* sdst, i32* H, u8* sel, i32* S, i32* pads, hipStream_t s) {
  const i64 nb = (M + SEL_BLK - 1) / SEL_BLK;
  k_sel_hist<<<(unsigned)nb, SEL_BLK, 0, s>>>(sk_idx, mv_dst, M, k, H, sdst);
  k_sel_scan<<<(unsigned)k, SELSCAN_BS, 0, s>>>(H, nb);
  if (k <= SELW_MAXK)
    k_sel_mark_w<<<(unsigned)nb, SEL_BLK, 0, s>>>(sdst, M, k, H, avail, sel, S, pads);
  else
    k_sel_mark<<<(unsigned)nb, SEL_BLK, 0, s>>>(sdst, M, k, H, avail, sel, S, pads);
}

void launch_commit(const i32* sk_idx, const u8* sel, i64 M, const i32* mv_v, const i32* mv_src,
                   const i32* mv_dst, i32* part, u8* part8, i32* pc, i32 k, i32* app,
                   i32* napp, hipStream_t s) {
  if (k <= MAXK_LDS)
    k_commit<true><<<grid_for(M, 256 * 16, 2048), 256, 0, s>>>(sk_idx, sel, M, mv_v, mv_src,
                                                               mv_dst, part, part8, pc, k, app,
                                                               napp);
  else
    k_commit<false><<<grid_for(M, 256, 1u << 16), 256, 0, s>>>(sk_idx, sel, M, mv_v, mv_src,
                                                               mv_dst, part, part8, pc, k, app,
                                                               napp);
}

void launch_gain_tallies(i64 n, const i32* part, const TallyRef& T, i32 k, const float* bg,
                         const u8* src_big, float* gain, i32* target, float* cand, int phase,
                         u64* cut_acc, int lanes, hipStream_t s) {
  const size_t lds = (size_t)2 * k * 4;
  // one lane per vertex on a level with byte rows (k <= 64); int rows are read
  // by 4 lanes per vertex unless the caller asks for another width
  if (T.t16 && !T.all_int && k <= TALLY16_MAXK) {
    const unsigned nb = grid_for(n, 256 / 4, 8192), nbb = grid_for(T.nbig, 256, 1024);
    k_gain_tallies16<4><<<nb + nbb, 256, 0, s>>>(n, part, T, k, bg, src_big, gain, target, cand,
                                                 phase, cut_acc, nb);
    return;
  }
  if (k <= GT1_MAXK && lanes == 1 && !T.all_int) {
    const unsigned nb = grid_for(n, 256, 8192), nbb = grid_for(T.nbig, 256, 1024);
    k_gain_tallies1<<<nb + nbb, 256, 0, s>>>(n, part, T, k, bg, src_big, gain, target, cand,
                                             phase, cut_acc, nb);
    return;
  }
  if (lanes == 1) lanes = 4;
  if (lanes == 2)
    k_gain_tallies<2, 8><<<grid_for(n, 128, 4096), 256, lds, s>>>(
        n, part, T, k, bg, src_big, gain, target, cand, phase, cut_acc);
  else if (lanes == 4)
    k_gain_tallies<4, 4><<<grid_for(n, 64, 4096), 256, lds, s>>>(
        n, part, T, k, bg, src_big, gain, target, cand, phase, cut_acc);
  else if (lanes == 8)
    k_gain_tallies<8, 2><<<grid_for(n, 32, 4096), 256, lds, s>>>(
        n, part, T, k, bg, src_big, gain, target, cand, phase, cut_acc);
  else
    k_gain_tallies<16, 1><<<grid_for(n, 16, 4096), 256, lds, s>>>(
        n, part, T, k, bg, src_big, gain, target, cand, phase, cut_acc);
}
void launch_tally_update(DGraph g, const TallyRef& T, i32 k, const i32* app, const i32* napp,
                         i64 max_app, const i32* mv_v, const i32* mv_src, const i32* mv_dst,
                         int unit, int skip_last, hipStream_t s) {
  k_tally_update<<<grid_for(max_app, 4, 4096), 256, 0, s>>>(g, T, k, app, napp, mv_v, mv_src,
                                                            mv_dst, unit, skip_last);
}
void launch_tally_cut(i64 n, const i32* part, const TallyRef& T, i32 k, u64* cut_acc,
                      hipStream_t s) {
  k_tally_cut<<<grid_for(n, 64, 4096), 256, 0, s>>>(n, part, T, k, cut_acc);
}
void launch_tally_last(DGraph g, i64 nedges, const i32* part, const TallyRef& T, i32 k, int unit,
                       hipStream_t s) {
  k_tally_last<<<1, 256, 0, s>>>(g, nedges, part, T, k, unit);
}
void launch_tally_class(DGraph g, i64 nedges, int unit, int all_int, i64 limit, i32* bix,
                        u32* bigmask, i32* nbig, hipStream_t s) {
  k_tally_class<<<grid_for(g.n, 256, 1u << 16), 256, 0, s>>>(g, nedges, unit, all_int, limit,
                                                              bix, bigmask, nbig);
}

void launch_big_list(const i32* bix, i64 n, i32* bigv, hipStream_t s) {
  if (n > 0) k_big_list<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(bix, n, bigv);
}

void launch_part8(const i32* part, i64 n, i32 k, u8* part8, hipStream_t s) {
  if (n > 0) k_part8<<<grid_for(n, 256, 1u << 16), 256, 0, s>>>(part, n, k, part8);
}

void launch_arc_source(DGraph g, i32* asrc, hipStream_t s) {
  k_arc_source<<<grid_for(g.n, 4, 1u << 16), 256, 0, s>>>(g, asrc);
}

void launch_pass_status(const i64* m, const u64* cut, const i32* ctr, const i32* pc, i32 k,
                        const i32* fit, i64* out, hipStream_t s) {
  k_pass_status<<<1, 256, 0, s>>>(m, cut, ctr, pc, k, fit, out);
}
void launch_acc_finish(u64* acc, u64* out, hipStream_t s) {
  k_acc_finish<<<1, ACC_SLOTS, 0, s>>>(acc, out);
}

void launch_cut_arcs(DGraph g, i64 e_end, const i32* asrc, const i32* part, const u8* part8,
                     i32 k, int unit, u64* acc, hipStream_t s) {
  const unsigned gr = grid_for(e_end, 256 * 8, 8192);
  if (part8 && k <= PART8_MAXK) {
    if (unit)
      k_cut_arcs<true, u8><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part8, k, acc);
    else
      k_cut_arcs<false, u8><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part8, k, acc);
  } else if (unit) {
    k_cut_arcs<true, i32><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part, k, acc);
  } else {
    k_cut_arcs<false, i32><<<gr, 256, 0, s>>>(asrc, g.adj, g.w, g.n, e_end, part, k, acc);
  }
}

i64 batch_coarsen_bytes(i64 maxn) {
  BeScratch X;
  return be_layout(maxn, true, &X, nullptr);
}
i64 batch_refine_bytes(i64 maxn) {
  BeScratch X;
  return be_layout(maxn, false, &X, nullptr);
}
i64 batch_record_bytes() { return (i64)sizeof(BeLevRec); }

void launch_batch_coarsen(const BatchGraph* graphs, const i32* order, i32 count,
                          const BatchBuffers& B, i64 budget, u64* stats, hipStream_t s) {
  const BePool pool{B.pool, B.pool_cap, B.pool_used};
  k_batch_coarsen<<<B.blocks_c, BE_T, 0, s>>>(graphs, order, count, B.next, B.scratch_c,
                                               B.maxn, batch_coarsen_bytes(B.maxn), pool,
                                               (BeLevRec*)B.recs, budget, stats);
}
template <int STILE>
void batch_refine_at(const BatchGraph* graphs, const i32* order, i32 count,
                     const BatchBuffers& B, u64* stats, hipStream_t s) {
  const i64 per = batch_refine_bytes(B.maxn);
  const BeLevRec* recs = (const BeLevRec*)B.recs;
  if (B.g == 2)
    k_batch_refine<STILE, 2><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 1,
                                                         B.scratch_r, B.maxn, per, recs, stats);
  else if (B.g == 8)
    k_batch_refine<STILE, 8><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 1,
                                                         B.scratch_r, B.maxn, per, recs, stats);
  else
    k_batch_refine<STILE, 4><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 1,
                                                         B.scratch_r, B.maxn, per, recs, stats);
}
void launch_batch_refine(const BatchGraph* graphs, const i32* order, i32 count,
                         const BatchBuffers& B, u64* stats, hipStream_t s) {
  if (B.stile == 2048)
    batch_refine_at<2048>(graphs, order, count, B, stats, s);
  else if (B.stile == 8192)
    batch_refine_at<8192>(graphs, order, count, B, stats, s);
  else
    batch_refine_at<4096>(graphs, order, count, B, stats, s);
}
void launch_batch_balance(const BatchGraph* graphs, const i32* order, i32 count,
                          const BatchBuffers& B, u64* stats, hipStream_t s) {
  const i64 per = batch_refine_bytes(B.maxn);
  if (B.g == 2)
    k_batch_balance<2><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 2, B.scratch_r,
                                                   B.maxn, per, stats);
  else if (B.g == 8)
    k_batch_balance<8><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 2, B.scratch_r,
                                                   B.maxn, per, stats);
  else
    k_batch_balance<4><<<B.blocks_r, BE_T, 0, s>>>(graphs, order, count, B.next + 2, B.scratch_r,
                                                   B.maxn, per, stats);
}
void launch_batch_engine(const BatchGraph* graphs, const i32* order, i32 count,
                         const BatchBuffers& B, i64 budget, u64* stats, hipStream_t s) {
  launch_batch_coarsen(graphs, order, count, B, budget, stats, s);
  launch_batch_refine(graphs, order, count, B, stats, s);
  if (B.balance) launch_batch_balance(graphs, order, count, B, stats, s);
}

}  // namespace tgp
