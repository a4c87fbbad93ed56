// tgp_driver.cpp -- host orchestration of the multilevel k-way partitioner on
// one GPU: the coarsening level loop, the initial-partition seam, projection
// and the two-phase refinement driver.  Control flow and every host-side fp32
// expression follow src/gpu_part.cu (iamrupamroyy/TheGraphPartitioning); all
// graph work runs in the gfx950 kernels of tgp_kernels.hip on one HIP stream.
//
// Built with -ffp-contract=off so the balance score keeps separately rounded
// multiply and add, as the reference's host code does.
#include "../../include/tgp.h"
#include "tgp_internal.h"
#include "../../include/tgp_learn.h"

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <atomic>
#include <algorithm>
#include <dlfcn.h>
#include <string>
#include <vector>

namespace tgp {

static void set_error(const std::string& s) { last_error() = s; }

#define HIP_TRY(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + __FILE__ + \
                ":" + std::to_string(__LINE__));                                    \
      throw Fail(e_ == hipErrorOutOfMemory ? TGP_ENOMEM : TGP_EHIP);                \
    }                                                                               \
  } while (0)

struct Fail {
  int code;
  explicit Fail(int c) : code(c) {}
};

// ---------------------------------------------------------------- device pool --
// Grow-only cache of device blocks keyed by rounded size; blocks return to the
// cache when released and are freed when the context is destroyed.
//
// Arena mode (set_arena): every block comes from one fixed device region the
// caller owns -- first fit over a list of free ranges, coalesced on release --
// and the pool never calls hipMalloc or hipFree.  Used by the entry point that
// allocates nothing (tgp_partition_noalloc).
class DevicePool {
 public:
  ~DevicePool() {
    for (auto& kv : free_)
      for (void* p : kv.second) (void)hipFree(p);
  }
  void set_arena(void* base, size_t cap) {
    arena_ = (char*)base;
    arena_cap_ = cap;
    ranges_.clear();
    ranges_[0] = cap;
  }
  bool arena() const { return arena_ != nullptr; }
  size_t arena_cap() const { return arena_cap_; }
  size_t arena_free() const {
    size_t t = 0;
    for (const auto& r : ranges_) t += r.second;
    return t;
  }
  static size_t round_size(size_t b) {
    if (b < 256) return 256;
    size_t p = 1;
    while (p < b) p <<= 1;
    const size_t step = p >= 2048 ? p / 8 : p;
    return ((b + step - 1) / step) * step;
  }
  void* get(size_t bytes) {
    if (arena_) return arena_get(bytes);
    const size_t r = round_size(bytes);
    auto it = free_.find(r);
    if (it != free_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, r);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of " + std::to_string(r) + " bytes failed");
      throw Fail(TGP_ENOMEM);
    }
    return p;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    if (arena_) {
      arena_put(p, bytes);
      return;
    }
    free_[round_size(bytes)].push_back(p);
  }
  // bytes held in released blocks (hipMemGetInfo counts them as used; get()
  // returns them to the runtime when an allocation would otherwise fail)
  size_t cached_bytes() const {
    size_t t = 0;
    for (const auto& kv : free_) t += kv.first * kv.second.size();
    return t;
  }
  void trim() {
    if (arena_) return;
    for (auto& kv : free_) {
      for (void* p : kv.second) (void)hipFree(p);
      kv.second.clear();
    }
  }

 private:
  static size_t arena_round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }
  void* arena_get(size_t bytes) {
    const size_t r = arena_round(bytes);
    for (auto it = ranges_.begin(); it != ranges_.end(); ++it) {
      if (it->second < r) continue;
      const size_t off = it->first, len = it->second;
      ranges_.erase(it);
      if (len > r) ranges_[off + r] = len - r;
      return arena_ + off;
    }
    set_error("the no-allocation entry point's scratch of " + std::to_string(arena_cap_ >> 20) +
              " MiB is exhausted (" + std::to_string(r) + " more bytes needed, " +
              std::to_string(arena_free()) + " free in pieces)");
    throw Fail(TGP_ENOMEM);
  }
  void arena_put(void* p, size_t bytes) {
    size_t off = (size_t)((char*)p - arena_), len = arena_round(bytes);
    auto next = ranges_.lower_bound(off);
    if (next != ranges_.end() && off + len == next->first) {
      len += next->second;
      next = ranges_.erase(next);
    }
    if (next != ranges_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    ranges_[off] = len;
  }

  std::map<size_t, std::vector<void*>> free_;
  char* arena_ = nullptr;
  size_t arena_cap_ = 0;
  std::map<size_t, size_t> ranges_;  // arena mode: free ranges, offset -> length
};

template <typename T>
struct DBuf {
  DevicePool* pool = nullptr;
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(DevicePool* pl, size_t cnt) : pool(pl), p((T*)pl->get(cnt * sizeof(T) + 16)), n(cnt) {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    release();
    pool = o.pool;
    p = o.p;
    n = o.n;
    o.p = nullptr;
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p && pool) pool->put(p, n * sizeof(T) + 16);
    p = nullptr;
  }
  void ensure(DevicePool* pl, size_t cnt) {
    if (p && n >= cnt) return;
    release();
    pool = pl;
    p = (T*)pl->get(cnt * sizeof(T) + 16);
    n = cnt;
  }
};

// ------------------------------------------------------------------ profiling --
struct Prof {
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> pool;
  struct Rec {
    int kid;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  double ms[KID_COUNT] = {};
  i64 launches[KID_COUNT] = {};
  double bytes[KID_COUNT] = {};
  int cur = -1;
  hipEvent_t cur_ev = nullptr;

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    return e;
  }
  void begin(int kid) {
    launches[kid]++;
    if (!on) return;
    cur = kid;
    cur_ev = ev();
    HIP_TRY(hipEventRecord(cur_ev, s));
  }
  void end(int kid, double algo_bytes) {
    bytes[kid] += algo_bytes;
    if (!on) return;
    hipEvent_t b = ev();
    HIP_TRY(hipEventRecord(b, s));
    pending.push_back(Rec{kid, cur_ev, b});
    cur = -1;
  }
  // takes the finished records (events complete in stream order, so the
  // unfinished ones are the last few: they stay pending)
  void flush() {
    if (!on || pending.empty()) return;
    size_t done = 0;
    for (; done < pending.size(); done++) {
      const Rec& r = pending[done];
      const hipError_t qa = hipEventQuery(r.a), qb = hipEventQuery(r.b);
      if (qa == hipErrorNotReady || qb == hipErrorNotReady) break;
      HIP_TRY(qa);
      HIP_TRY(qb);
      float t = 0;
      // the runtime may still report a finished pair as not ready: keep it
      const hipError_t e = hipEventElapsedTime(&t, r.a, r.b);
      if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        break;
      }
      HIP_TRY(e);
      ms[r.kid] += t;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.erase(pending.begin(), pending.begin() + (std::ptrdiff_t)done);
  }
  ~Prof() {
    for (auto& r : pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// ---------------------------------------------------------------------- context --
// Per-level data derived once for the refinement kernels (valid for one value
// of the caller's nedges, which only differs from m on the coarsest level).
struct LevelAux {
  bool valid = false;
  i64 nedges = -1;
  i64 e_end = 0;      // arcs [0, e_end) = the reference's row ranges
  i64 last_begin = 0; // xadj[n-1]
  i64 nheavy = 0;     // rows longer than HEAVY_DEG
  int unit = 0;       // all weights are 1
  DBuf<i32> asrc;     // arc -> source vertex (built on first use: arc_source())
  bool asrc_ok = false;
  DBuf<i32> heavy;    // heavy rows
  DBuf<i32> mid, medium, huge;  // heavy rows split at 512 and 4096 arcs
  i64 nmid = 0, nmedium = 0, nhuge = 0;
  DBuf<i32> ch_row;          // chunks of the long rows: row index in `huge`, arc range
  DBuf<i64> ch_beg, ch_end;
  i64 nch = 0;
  DBuf<i32> gtally;          // nhuge * k tallies and nhuge |w| sums, zero between passes
  DBuf<i64> gabs;
  int all_exact = 0;         // every row's sum of |w| is at most 2^24
  DBuf<i32> over;            // [0] = count, then the rows whose sum of |w| passes 2^24
  i64 nover = 0;
  DBuf<u32> ctr;             // work-item counter of the fused gain kernel
  GainLists lists() const {
    GainLists L;
    L.mid = mid.p;
    L.nmid = nmid;
    L.medium = medium.p;
    L.nmedium = nmedium;
    L.huge = huge.p;
    L.nhuge = nhuge;
    L.ch_row = ch_row.p;
    L.ch_beg = ch_beg.p;
    L.ch_end = ch_end.p;
    L.nch = nch;
    L.gtally = gtally.p;
    L.gabs = gabs.p;
    L.all_exact = all_exact;
    L.ctr = ctr.p;
    return L;
  }
};

struct Level {
  i64 n = 0, m = 0;
  bool symmetric = false;  // every arc s->x (s != x) has its reverse
  bool simple_ws = false;  // and every row is simple and each reverse has the same weight
  bool sorted = false;     // every row strictly ascending (checked on level 0 only)
  bool unitw = false;      // every arc weight is 1 (checked on level 0 only)
  bool examined = false;   // the symmetry check ran on this level
  LevelAux aux;
  DBuf<i64> xadj;
  DBuf<i32> adj, w;
  DBuf<i32> tb;  // 2n: coarse vertex -> its fine vertices on the level above
  const i64* xadj_p = nullptr;
  const i32* adj_p = nullptr;
  const i32* w_p = nullptr;
  DGraph dg() const { return DGraph{n, m, xadj_p, adj_p, w_p}; }
};

struct Context {
  int device = 0;
  hipStream_t s = nullptr;
  DevicePool pool;
  i64* pinned = nullptr;  // small host staging area
  static constexpr size_t PINNED_ELEMS = 8192 + 64;
  Prof prof;
  i64 clock_khz = 0;  // device wall-clock rate

  // arena != nullptr: the pool works inside that device region and the host
  // staging area is `host_stage` (ordinary memory), so the context allocates
  // nothing after construction
  explicit Context(void* arena = nullptr, size_t arena_bytes = 0, i64* host_stage = nullptr) {
    HIP_TRY(hipGetDevice(&device));
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    clock_khz = std::max(khz, 1000);
    if (arena) {
      pool.set_arena(arena, arena_bytes);
      pinned = host_stage;
    } else {
      HIP_TRY(hipHostMalloc((void**)&pinned, PINNED_ELEMS * sizeof(i64), hipHostMallocDefault));
      own_pinned = true;
    }
    prof.s = s;
    HIP_TRY(hipEventCreateWithFlags(&mark_ev, hipEventDisableTiming));
    HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_side, hipEventDisableTiming));
  }
  ~Context() {
    if (pinned && own_pinned) (void)hipHostFree(pinned);
    if (mark_ev) (void)hipEventDestroy(mark_ev);
    if (ev_main) (void)hipEventDestroy(ev_main);
    if (ev_side) (void)hipEventDestroy(ev_side);
    if (s2) (void)hipStreamDestroy(s2);
    if (s) (void)hipStreamDestroy(s);
  }
  bool own_pinned = false;
  void sync() {
    HIP_TRY(hipStreamSynchronize(s));
    prof.flush();
  }
  // mark(): the work enqueued so far; wait_mark(): wait for it only (work
  // enqueued after the mark keeps running)
  hipEvent_t mark_ev = nullptr;
  // a second stream (default priority) for work that overlaps the main one:
  // ev_main orders it after the main stream's work so far, ev_side marks its end
  hipStream_t s2 = nullptr;
  hipEvent_t ev_main = nullptr, ev_side = nullptr;
  void mark() { HIP_TRY(hipEventRecord(mark_ev, s)); }
  void wait_mark() {
    HIP_TRY(hipEventSynchronize(mark_ev));
    prof.flush();
  }
};

// ------------------------------------------------------------------------ run --
struct Tracer {
  tgp_trace* buf = nullptr;
  i64 cap = 0, len = 0;
  i64* len_out = nullptr;  // mirrors len into tgp_stats so failures keep their trace
  void push(int kind, int level, int phase, int pass, i64 a, i64 b, i64 c, i64 d, u64 h) {
    if (buf && len < cap) {
      tgp_trace& t = buf[len];
      t.kind = kind;
      t.level = level;
      t.phase = phase;
      t.pass = pass;
      t.a = a;
      t.b = b;
      t.c = c;
      t.d = d;
      t.hash = h;
    }
    len++;
    if (len_out) *len_out = len;
  }
};

static u64 fnv_host(const i32* v, i64 n) {
  u64 h = 1469598103934665603ull;
  for (i64 i = 0; i < n; i++) {
    const u32 x = (u32)v[i];
    for (int b = 0; b < 4; b++) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

class Run {
 public:
  Run(Context& c, const tgp_options& o, tgp_stats* st) : C(c), O(o), S(st) {
    k = o.k;
    wide = o.wide != 0;
    repair = o.repair != 0;
    if (st) {
      tr.buf = st->trace;
      tr.cap = st->trace_cap;
      tr.len_out = &st->trace_len;
    }
  }

  // the symmetry check may still run on the second stream when a run ends
  // early (an error): its buffers are not handed back before it finishes
  ~Run() {
    if (sp_active) (void)hipStreamSynchronize(C.s2);
  }

  int go(const tgp_graph* g, i32* part_out, i32* pc_out, i64* cut_out);
  void setup_level0(const tgp_graph* g);
  // the balance stage alone on a given partition of g (tgp_balance)
  int balance_only(const tgp_graph* g, i32* part_io, i32* pc_out, i64* cut_out);

 private:
  Context& C;
  const tgp_options& O;
  tgp_stats* S;
  Tracer tr;
  i32 k = 0;
  bool wide = false, repair = false;
  int passes_total = 0;
  i64 sorts_skipped = 0;  // passes whose move sort could not change the selection
  i64 match_rounds = 0;
  double initpart_s = 0;

  std::vector<std::unique_ptr<Level>> L;
  std::vector<i64> ecount;

  // shared scratch sized for the finest level
  DBuf<i64> scan_sc;
  DBuf<i64> small;  // device scalars: [0] total, [1] cut, [2..] misc
  DBuf<u64> accs;   // accumulator slots (ACC_ELEMS), zero between uses
  DBuf<i32> small32;

  hipStream_t st() { return C.s; }
  ScanScratch scratch() { return ScanScratch{scan_sc.p, (i64)scan_sc.n}; }

  u64 hash_dev(const i32* d, i64 n) {
    std::vector<i32> h((size_t)n);
    HIP_TRY(hipMemcpyAsync(h.data(), d, n * sizeof(i32), hipMemcpyDeviceToHost, st()));
    C.sync();
    return fnv_host(h.data(), n);
  }
  i64 read_i64(const i64* d) {
    HIP_TRY(hipMemcpyAsync(C.pinned, d, sizeof(i64), hipMemcpyDeviceToHost, st()));
    C.sync();
    return C.pinned[0];
  }

  i64 wrap_cut(u64 acc) const {
    if (!wide) {
      const i32 c = (i32)(u32)acc;
      return (i64)(c / 2);
    }
    return (i64)acc / 2;
  }

  void coarsen();
  void match_level(const Level& lv, DBuf<i32>& mate);
  // overlap: on strictly ascending rows, run the search on the second stream
  // and take the level as symmetric until join_symmetric (returns true then)
  bool check_symmetric(Level& lv, bool overlap = false);
  // the overlapped check's verdict into lv; false when the level turned out
  // not symmetric (whatever was derived from the assumption must be redone)
  bool join_symmetric(Level& lv);
  DBuf<i32> sp_flags;
  DBuf<u64> sp_cnt;
  DBuf<u8> sp_cls;
  DBuf<i32> sp_crow;
  bool sp_active = false;
  int initial_partition(DBuf<i32>& part, i64* npairs);
  i64 compute_cut(Level& lv, i64 nedges, const i32* part);
  const i32* arc_source(Level& lv);
  LevelAux& aux(Level& lv, i64 nedges);
  void refine(int level, i64 nedges, DBuf<i32>& part, DBuf<i32>& pc, i64* cut,
              bool cut_known = true);
  // cap > 0: capacity cap - count per part instead of the reference's rule
  void learned_passes(int level, i64 nedges, DBuf<i32>& part, DBuf<i32>& pc, i64 cap = 0);
  void balance_finest(DBuf<i32>& part, DBuf<i32>& pc);
  // the refinement trace callback with a host copy of part
  void dump_part(int level, int phase, int pass, const i32* part, i64 n) {
    std::vector<i32> h((size_t)n);
    HIP_TRY(hipMemcpyAsync(h.data(), part, n * sizeof(i32), hipMemcpyDeviceToHost, C.s));
    C.sync();
    O.pass_dump(level, phase, pass, n, h.data(), O.pass_dump_user);
  }
};

static const i64 HEAVY_DEG = 64;
// most rows with a sum of |w| over 2^24 that a level may have and still keep
// part tallies (their gains are gathered again on every pass)
static const i64 OVER_ROWS_MAX = 65536;

static i64 env_i64(const char* name, i64 dflt) {
  const char* v = std::getenv(name);
  return (v && *v) ? std::atoll(v) : dflt;
}

// Lanes per light row in the token pass (0 = one arc stream per 64 rows,
// 1, 4 or 8: TGP_TOKEN_LANES) and per vertex in the gain pass over kept
// tallies (1 when k <= 64, else 4 by default; 2, 4, 8 or 16: TGP_GAIN_LANES).
static int token_lanes() { return (int)env_i64("TGP_TOKEN_LANES", 0); }
static int gain_lanes() { return (int)env_i64("TGP_GAIN_LANES", 1); }

void Run::match_level(const Level& lv, DBuf<i32>& mate) {
  const i64 n = lv.n;
  static const i64 MATCH_WINDOW = std::max<i64>(64, env_i64("TGP_MATCH_WINDOW", 16384));
  static const bool trace_match = env_i64("TGP_MATCH_TRACE", 0) != 0;
  // Tuning and test knobs, read at every call: TGP_MATCH_CHAIN=0 selects the
  // round-based kernels on symmetric levels; TGP_CHAIN_BLOCKS sets the chained
  // kernel's grid, TGP_CHAIN_SG and TGP_CHAIN_PER its lanes per vertex and
  // row entries per lane;
  // TGP_CHAIN_BUDGET_US bounds one chained launch (100 ms).
  const bool MATCH_CHAIN = env_i64("TGP_MATCH_CHAIN", 1) != 0;
  const int CHAIN_BLOCKS = (int)env_i64("TGP_CHAIN_BLOCKS", 0);  // 0: all resident blocks
  const int CHAIN_PER = (int)env_i64("TGP_CHAIN_PER", 8);
  // Lanes per vertex: 8 when every row fits 8 lanes' registers (8 vertices
  // per wave, so 4x as many vertices in flight: on a 3-D grid the decisions
  // run as wavefronts through consecutive ids and the number in flight sets
  // the pace, 1.54 -> 0.50 s on the 256^3 grid) or the level has at least
  // 2^22 vertices (with the long rows' best-entry lists, R-MAT 24 ef16 k=64
  // 0.796 -> 0.781 s, weighted 0.802 -> 0.791 s; R-MAT 22 even, R-MAT 20
  // 1.5 % slower), else 16 (rows longer than 128 arcs are scanned by the
  // whole wave; 16 against 32 lanes: R-MAT 24 ef16 k=64 1.366 -> 1.312 s,
  // R-MAT 22 0.431 -> 0.420 s).
  int CHAIN_SG = (int)env_i64("TGP_CHAIN_SG", 0);
  if (CHAIN_SG <= 0 && lv.symmetric && MATCH_CHAIN) {
    DBuf<u64> mr(&C.pool, 1);
    HIP_TRY(hipMemsetAsync(mr.p, 0, sizeof(u64), st()));
    launch_max_row(lv.dg(), mr.p, st());
    HIP_TRY(hipMemcpyAsync(C.pinned, mr.p, sizeof(u64), hipMemcpyDeviceToHost, st()));
    C.sync();
    CHAIN_SG = ((i64)C.pinned[0] <= (i64)8 * CHAIN_PER || n >= ((i64)1 << 22)) ? 8 : 16;
  }
  if (CHAIN_SG <= 0) CHAIN_SG = 32;
  const i64 chain_budget =
      std::max<i64>(1, C.clock_khz * env_i64("TGP_CHAIN_BUDGET_US", 100000) / 1000);
  const auto t_start = std::chrono::steady_clock::now();
  DBuf<u8> dec(&C.pool, n);
  DBuf<i32> s1(&C.pool, n), prop(&C.pool, n);
  // [0] window low end, [1..2] scan starts by round parity, [3] first-undecided
  // accumulator (n between uses)
  DBuf<i64> lo(&C.pool, 4);
  DBuf<u64> R;
  DBuf<i32> s2;
  if (lv.symmetric) s2 = DBuf<i32>(&C.pool, n);
  else R = DBuf<u64>(&C.pool, n);
  DBuf<i32> decision, req, nreq;
  if (lv.symmetric) {
    decision = DBuf<i32>(&C.pool, n);
    req = DBuf<i32>(&C.pool, 2 * 3 * std::min<i64>(MATCH_WINDOW, n) + 2);
    nreq = DBuf<i32>(&C.pool, 2);
    HIP_TRY(hipMemsetAsync(nreq.p, 0, 2 * sizeof(i32), st()));
  }
  // algorithmic bytes of one level's matching: offsets, arcs and weights, state
  const double match_bytes = (double)(lv.n + 1) * 8 + 12.0 * lv.m + 4.0 * lv.n;
  C.prof.begin(KID_MATCH);
  if (lv.symmetric)
    launch_match_sym_init(n, mate.p, dec.p, s1.p, s2.p, prop.p, lo.p, st());
  else
    launch_match_init(n, mate.p, dec.p, s1.p, R.p, lo.p, st());
  u32 round = 0;
  const DGraph g = lv.dg();
  if (lv.symmetric && MATCH_CHAIN) {
    // Chained matching; each launch is bounded by a time budget.  Should
    // launches stop advancing the first undecided vertex, the round-based
    // kernels below finish the level from the same (consistent) state.
    DBuf<u32> ticket(&C.pool, 1);
    DBuf<i32> stv(&C.pool, n);
    DBuf<u64> cstats;
    if (trace_match) {
      cstats = DBuf<u64>(&C.pool, 12);
      HIP_TRY(hipMemsetAsync(cstats.p, 0, 12 * sizeof(u64), st()));
    }
    launch_fill_i32(stv.p, n, -1, st());
    HIP_TRY(hipMemsetAsync(ticket.p, 0, sizeof(u32), st()));
    // Lower-neighbour lists of the rows the kernel scans with the whole wave
    // (longer than CHAIN_SG * CHAIN_PER): "no undecided neighbour below t"
    // then reads only the entries below t instead of the whole row; on a
    // power-law level the hub rows are long and their lower neighbours few.
    // TGP_CHAIN_LOW=0 scans whole rows.
    DBuf<i64> lowcnt, lowoff;
    DBuf<i32> lowadj;
    ChainLow low{nullptr, nullptr};
    if (env_i64("TGP_CHAIN_LOW", 1) != 0) {
      const i64 minlen = (i64)CHAIN_SG * CHAIN_PER;
      lowcnt = DBuf<i64>(&C.pool, n);
      lowoff = DBuf<i64>(&C.pool, n + 1);
      DBuf<i32> llist(&C.pool, n), nll(&C.pool, 1);
      launch_fill_i64(lowcnt.p, n, 0, st());
      HIP_TRY(hipMemsetAsync(nll.p, 0, sizeof(i32), st()));
      launch_chain_low_count(g, minlen, llist.p, nll.p, lowcnt.p, st());
      launch_scan_excl(lowcnt.p, lowoff.p, n, lowoff.p + n, scratch(), st());
      const i64 tot = read_i64(lowoff.p + n);
      lowadj = DBuf<i32>(&C.pool, std::max<i64>(tot, 1));
      if (tot > 0) launch_chain_low_write(g, llist.p, nll.p, lowoff.p, lowadj.p, st());
      low = ChainLow{lowoff.p, lowadj.p};
    }
    // Best-entry lists of the rows longer than TGP_CHAIN_LIST_MIN (4096) arcs
    // on weighted levels (TGP_CHAIN_LISTS=0: none): a long row's proposal is
    // then its list's first free entry, and the row is scanned only once the
    // list is used up.
    DBuf<i32> lidx, lrows, lnl, llists, lcnt;
    if (!lv.unitw && env_i64("TGP_CHAIN_LISTS", 1) != 0) {
      const i64 lmin =
          std::max<i64>(env_i64("TGP_CHAIN_LIST_MIN", 4096), (i64)CHAIN_SG * CHAIN_PER);
      const i64 maxl = lv.m / lmin + 1;
      lidx = DBuf<i32>(&C.pool, n);
      lrows = DBuf<i32>(&C.pool, maxl);
      lnl = DBuf<i32>(&C.pool, 1);
      llists = DBuf<i32>(&C.pool, maxl * CHAIN_LIST_LEN);
      lcnt = DBuf<i32>(&C.pool, maxl);
      HIP_TRY(hipMemsetAsync(lnl.p, 0, sizeof(i32), st()));
      launch_chain_lists(g, lmin, lidx.p, lrows.p, lnl.p, llists.p, lcnt.p, st());
      low.lidx = lidx.p;
      low.lists = llists.p;
      low.lcnt = lcnt.p;
    }
    // TGP_CHAIN_EXACT_SCANS=0 keeps the full scans on sorted / unit-weight levels
    if (env_i64("TGP_CHAIN_EXACT_SCANS", 1) != 0) {
      low.sorted = lv.sorted ? 1 : 0;
      low.unitw = lv.unitw ? 1 : 0;
    }
    C.prof.end(KID_MATCH, 0);
    i64 prev = -1;
    int stuck = 0;
    while (stuck < 3) {
      C.prof.begin(KID_CHAIN);
      launch_match_chain(g, low, stv.p, s1.p, s2.p, prop.p, lo.p, ticket.p, chain_budget,
                         CHAIN_BLOCKS, CHAIN_SG, CHAIN_PER, cstats.p, st());
      C.prof.end(KID_CHAIN, round == 0 ? match_bytes : 0.0);
      round++;
      launch_first_undecided(stv.p, n, lo.p, st());
      HIP_TRY(hipMemsetAsync(ticket.p, 0, sizeof(u32), st()));
      HIP_TRY(hipMemcpyAsync(C.pinned, lo.p, sizeof(i64), hipMemcpyDeviceToHost, st()));
      C.sync();
      if (trace_match)
        std::fprintf(stderr, "#   launch %u: first undecided %lld of %lld at %.2f ms\n", round,
                     (long long)C.pinned[0], (long long)n,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                               t_start).count());
      if (C.pinned[0] >= n) break;
      stuck = C.pinned[0] > prev ? 0 : stuck + 1;
      prev = C.pinned[0];
    }
    C.prof.begin(KID_MATCH);
    launch_chain_export(stv.p, n, mate.p, dec.p, st());
    if (C.pinned[0] >= n) {
      C.prof.end(KID_MATCH, 0);
      match_rounds += round;
      if (trace_match) {
        u64 cs[12];
        HIP_TRY(hipMemcpy(cs, cstats.p, sizeof(cs), hipMemcpyDeviceToHost));
        std::fprintf(stderr,
                     "# match n=%lld m=%lld chained launches=%u ms=%.2f iter=%llu idle=%llu "
                     "apS=%llu aL=%llu pL=%llu fetch=%llu cS=%llu cL=%llu dec=%llu arcsA=%llu "
                     "arcsP=%llu arcsC=%llu\n",
                     (long long)n, (long long)lv.m, round,
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                               t_start).count(),
                     (unsigned long long)cs[0], (unsigned long long)cs[1],
                     (unsigned long long)cs[2], (unsigned long long)cs[3],
                     (unsigned long long)cs[4], (unsigned long long)cs[5],
                     (unsigned long long)cs[6], (unsigned long long)cs[7],
                     (unsigned long long)cs[8], (unsigned long long)cs[9],
                     (unsigned long long)cs[10], (unsigned long long)cs[11]);
      }
      return;
    }
    // scan starts of the round kernels: both parity slots at the first undecided vertex
    launch_fill_i64(lo.p + 1, 2, C.pinned[0], st());
  }
  // Rounds are issued in batches; the window's low end is read back between them.
  int batch = 4;
  while (true) {
    for (int b = 0; b < batch; b++) {
      round++;
      if (lv.symmetric)
        launch_match_sym_round(g, mate.p, dec.p, s1.p, s2.p, prop.p, decision.p, lo.p,
                               MATCH_WINDOW, req.p, nreq.p, round, st());
      else
        launch_match_round(g, mate.p, dec.p, s1.p, R.p, prop.p, lo.p, MATCH_WINDOW, round, st());
    }
    HIP_TRY(hipMemcpyAsync(C.pinned, lo.p, sizeof(i64), hipMemcpyDeviceToHost, st()));
    HIP_TRY(hipStreamSynchronize(st()));
    if (C.pinned[0] >= n) break;
    // Each round decides at least the smallest undecided vertex, so more
    // rounds than vertices means a broken invariant: stop instead of spinning.
    if ((i64)round > n + 2 * 64) {
      set_error("matching made no progress");
      throw Fail(TGP_EHIP);
    }
    if (batch < 64) batch *= 2;
  }
  C.prof.end(KID_MATCH, match_bytes);
  match_rounds += round;
  if (trace_match)
    std::fprintf(stderr, "# match n=%lld m=%lld symmetric=%d rounds=%u ms=%.2f\n", (long long)n,
                 (long long)lv.m, (int)lv.symmetric, round,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                           t_start).count());
}

bool Run::check_symmetric(Level& lv, bool overlap) {
  lv.examined = true;
  DBuf<i32> flags(&C.pool, 6);
  DBuf<u64> cnt(&C.pool, 2);
  HIP_TRY(hipMemsetAsync(flags.p, 0, 6 * sizeof(i32), st()));
  HIP_TRY(hipMemsetAsync(cnt.p, 0, 2 * sizeof(u64), st()));
  // the row of each arc chunk's first arc, for both checks
  DBuf<i32> crow(&C.pool, std::max<i64>(sym_check_chunks(lv.m), 1));
  C.prof.begin(KID_MISC);
  launch_rows_strict(lv.dg(), lv.m, crow.p, flags.p, st());
  C.prof.end(KID_MISC, 8.0 * lv.m);
  HIP_TRY(hipMemcpyAsync(C.pinned, flags.p, 6 * sizeof(i32), hipMemcpyDeviceToHost, st()));
  C.sync();
  const bool strict = ((const i32*)C.pinned)[3] == 0;
  const bool nonunit = ((const i32*)C.pinned)[5] != 0;
  lv.sorted = strict;
  if (strict && overlap) {
    // the search runs beside the level's matching, which assumes symmetry
    // (and, on a symmetric level, the weights rows_strict saw decide unitw)
    sp_flags = std::move(flags);
    sp_cnt = std::move(cnt);
    sp_cls = DBuf<u8>(&C.pool, std::max<i64>(lv.n, 1));
    sp_crow = std::move(crow);
    HIP_TRY(hipEventRecord(C.ev_main, C.s));
    HIP_TRY(hipStreamWaitEvent(C.s2, C.ev_main, 0));
    launch_sym_check_sorted(lv.dg(), lv.m, sp_cls.p, sp_crow.p, sp_flags.p, sp_cnt.p, C.s2);
    HIP_TRY(hipEventRecord(C.ev_side, C.s2));
    sp_active = true;
    lv.symmetric = true;
    lv.simple_ws = false;
    lv.unitw = !nonunit;
    return true;
  }
  // strictly ascending rows hold no duplicates: each pair's searching arc
  // finds its reverse by binary search in the sorted row; other rows go
  // through a hash table of every arc
  DBuf<u64> table;
  DBuf<i32> wt;
  C.prof.begin(KID_MISC);
  DBuf<u8> cls;
  if (strict) {
    cls = DBuf<u8>(&C.pool, std::max<i64>(lv.n, 1));
    launch_sym_check_sorted(lv.dg(), lv.m, cls.p, crow.p, flags.p, cnt.p, st());
  } else {
    i64 cap = 1;
    while (cap < 2 * lv.m + 2) cap <<= 1;
    table = DBuf<u64>(&C.pool, cap);
    wt = DBuf<i32>(&C.pool, cap);
    launch_sym_check(lv.dg(), table.p, wt.p, cap, flags.p, st());
  }
  C.prof.end(KID_MISC, (strict ? 14.0 : 28.0) * lv.m);
  HIP_TRY(hipMemcpyAsync(C.pinned, flags.p, 5 * sizeof(i32), hipMemcpyDeviceToHost, st()));
  HIP_TRY(hipMemcpyAsync(C.pinned + 3, cnt.p, 2 * sizeof(u64), hipMemcpyDeviceToHost, st()));
  C.sync();
  const i32* f = (const i32*)C.pinned;
  const bool sym = f[0] == 0 && (!strict || C.pinned[3] == C.pinned[4]);
  lv.symmetric = sym;
  lv.simple_ws = sym && f[1] == 0 && f[2] == 0;
  // unit weights are read off the sorted check's found pairs (all arcs of a
  // symmetric level between distinct vertices in range)
  lv.unitw = sym && strict && f[4] == 0;
  return false;
}

bool Run::join_symmetric(Level& lv) {
  if (!sp_active) return true;
  HIP_TRY(hipEventSynchronize(C.ev_side));
  sp_active = false;
  HIP_TRY(hipMemcpyAsync(C.pinned, sp_flags.p, 5 * sizeof(i32), hipMemcpyDeviceToHost, st()));
  HIP_TRY(hipMemcpyAsync(C.pinned + 3, sp_cnt.p, 2 * sizeof(u64), hipMemcpyDeviceToHost, st()));
  C.sync();
  const i32* f = (const i32*)C.pinned;
  const bool sym = f[0] == 0 && C.pinned[3] == C.pinned[4];
  lv.symmetric = sym;
  lv.simple_ws = sym && f[1] == 0 && f[2] == 0;
  lv.unitw = sym && f[4] == 0;
  sp_flags = DBuf<i32>();
  sp_cnt = DBuf<u64>();
  sp_cls = DBuf<u8>();
  sp_crow = DBuf<i32>();
  return sym;
}

void Run::coarsen() {
  const i64 n0 = L[0]->n;
  const i64 target = O.coarsen_floor > 0 ? O.coarsen_floor : std::max<i64>(500, n0 / 500);
  i64 ncur = n0;
  // On strictly ascending rows the symmetry search of level 0 runs on the
  // second stream beside the level's matching, which assumes symmetry; a
  // level that turns out not symmetric is matched again by the round kernels
  // (TGP_SYM_OVERLAP=0: check first; never in the no-allocation context,
  // whose staging memory is not pinned)
  const bool overlap = !C.pool.arena() && env_i64("TGP_SYM_OVERLAP", 1) != 0;
  bool spec = false;
  if (ncur > target) spec = check_symmetric(*L[0], overlap);
  DBuf<i32> mate(&C.pool, n0), new_ids(&C.pool, n0);
  DBuf<i64> flags(&C.pool, n0), pos(&C.pool, n0);
  while (ncur > target) {  // gpu_part.cu:1003
    // repaired mode: a contraction that kept more than 93 % of the vertices
    // ends the coarsening before the next matching (whose level would keep
    // more than 95 % on power-law graphs and be dropped by the stall stop)
    if (repair && L.size() >= 2 && 100 * L.back()->n > 93 * L[L.size() - 2]->n) break;
    Level& lv = *L.back();
    match_level(lv, mate);
    if (spec) {
      spec = false;
      if (!join_symmetric(lv)) match_level(lv, mate);
    }
    auto tb = DBuf<i32>(&C.pool, 2 * ncur);
    C.prof.begin(KID_RELABEL);
    launch_relabel(mate.p, ncur, flags.p, pos.p, small.p, new_ids.p, tb.p, scratch(), st());
    C.prof.end(KID_RELABEL, 16.0 * ncur + 8.0 * ncur);
    const i64 n_c = read_i64(small.p);
    if (n_c <= 0 || n_c >= ncur) break;                        // gpu_part.cu:1069
    if (repair && 20 * n_c > 19 * ncur) break;                 // repaired mode: stall stop
    std::unique_ptr<Level> c(new Level());
    c->n = n_c;
    c->xadj = DBuf<i64>(&C.pool, n_c + 1);
    DBuf<i64> cnt(&C.pool, n_c + 1);
    const DGraph g = lv.dg();
    i64 m_c = 0;
    C.prof.begin(KID_CONTRACT);
    if (!repair) {
      DBuf<i32> sadj(&C.pool, n_c * 64), sw(&C.pool, n_c * 64);
      launch_contract_ref(g, mate.p, new_ids.p, n_c, sadj.p, sw.p, cnt.p, st());
      launch_scan_excl(cnt.p, c->xadj.p, n_c, c->xadj.p + n_c, scratch(), st());
      m_c = read_i64(c->xadj.p + n_c);
      c->adj = DBuf<i32>(&C.pool, std::max<i64>(m_c, 1));
      c->w = DBuf<i32>(&C.pool, std::max<i64>(m_c, 1));
      if (m_c > 0) launch_compact_rows(sadj.p, sw.p, c->xadj.p, n_c, c->adj.p, c->w.p, st());
    } else {
      DBuf<i64> dsz(&C.pool, n_c + 1), dlong(&C.pool, n_c + 1), csz(&C.pool, n_c + 1),
          eoff(&C.pool, n_c + 1), toff(&C.pool, n_c + 1), nchk(&C.pool, n_c + 1),
          cbase(&C.pool, n_c + 1);
      DBuf<i32> med(&C.pool, n_c), meds(&C.pool, n_c), lng(&C.pool, n_c), ncls(&C.pool, 3);
      HIP_TRY(hipMemsetAsync(ncls.p, 0, 3 * sizeof(i32), st()));
      launch_rep_classify(lv.xadj_p, tb.p, n_c, dsz.p, dlong.p, csz.p, med.p, ncls.p, lng.p,
                          ncls.p + 1, nchk.p, meds.p, ncls.p + 2, st());
      launch_scan_excl(dlong.p, eoff.p, n_c, eoff.p + n_c, scratch(), st());
      launch_scan_excl(csz.p, toff.p, n_c, toff.p + n_c, scratch(), st());
      HIP_TRY(hipMemcpyAsync(C.pinned, eoff.p + n_c, sizeof(i64), hipMemcpyDeviceToHost, st()));
      HIP_TRY(hipMemcpyAsync(C.pinned + 1, toff.p + n_c, sizeof(i64), hipMemcpyDeviceToHost,
                             st()));
      HIP_TRY(hipMemcpyAsync(C.pinned + 2, ncls.p, 3 * sizeof(i32), hipMemcpyDeviceToHost,
                             st()));
      C.sync();
      const i64 etot = C.pinned[0], ttot = C.pinned[1];
      const i64 nmed = ((const i32*)(C.pinned + 2))[0], nlng = ((const i32*)(C.pinned + 2))[1];
      const i64 nmeds = ((const i32*)(C.pinned + 2))[2];
      RepLong RL;
      DBuf<u64> T;
      DBuf<i32> W, E, ch_row;
      DBuf<i64> ch_beg, chunk_cnt, choff;
      // bucketed long rows (default): TGP_REP_BUCKETS=0 takes the global tables
      // instead; TGP_REP_BUCKET_FILL lowers the per-bucket key limit and
      // TGP_REP_BUCKET_TARGET sets the entries per bucket aimed at (tests)
      RepBuckets RB;
      bool bucketed = false;
      DBuf<i32> lgb, brow, over;
      DBuf<i64> nbk, bbase, boff;
      DBuf<u64> bcnt, R, BK;
      DBuf<u32> bcur;
      DBuf<i32> BW;
      // entry words of the rows past 64 entries (medium and long)
      if (etot > 0) R = DBuf<u64>(&C.pool, etot);
      if (nlng > 0) {
        launch_scan_excl(nchk.p, cbase.p, nlng, cbase.p + nlng, scratch(), st());
        const i64 nch = read_i64(cbase.p + nlng);
        ch_row = DBuf<i32>(&C.pool, nch);
        ch_beg = DBuf<i64>(&C.pool, nch);
        chunk_cnt = DBuf<i64>(&C.pool, nch + 1);
        choff = DBuf<i64>(&C.pool, nch + 1);
        launch_rep_chunks(lng.p, nlng, dsz.p, cbase.p, ch_row.p, ch_beg.p, st());
        RL.lng = lng.p;
        RL.nl = nlng;
        RL.dsz = dsz.p;
        RL.cbase = cbase.p;
        RL.ch_row = ch_row.p;
        RL.ch_beg = ch_beg.p;
        RL.nch = nch;
        RL.eoff = eoff.p;
        RL.chunk_cnt = chunk_cnt.p;
        RL.choff = choff.p;
        if (env_i64("TGP_REP_BUCKETS", 1) != 0) {
          lgb = DBuf<i32>(&C.pool, nlng);
          nbk = DBuf<i64>(&C.pool, nlng);
          bbase = DBuf<i64>(&C.pool, nlng + 1);
          launch_rl_buckets(RL, env_i64("TGP_REP_BUCKET_TARGET", 0), lgb.p, nbk.p, st());
          launch_scan_excl(nbk.p, bbase.p, nlng, bbase.p + nlng, scratch(), st());
          const i64 nbt = read_i64(bbase.p + nlng);
          brow = DBuf<i32>(&C.pool, nbt);
          bcnt = DBuf<u64>(&C.pool, nbt);
          boff = DBuf<i64>(&C.pool, nbt + 1);
          bcur = DBuf<u32>(&C.pool, nbt);
          over = DBuf<i32>(&C.pool, 1);
          BK = DBuf<u64>(&C.pool, etot);
          BW = DBuf<i32>(&C.pool, etot);
          RB.lgb = lgb.p;
          RB.nbk = nbk.p;
          RB.bbase = bbase.p;
          RB.nbt = nbt;
          RB.brow = brow.p;
          RB.bcnt = bcnt.p;
          RB.boff = boff.p;
          RB.bcur = bcur.p;
          RB.R = R.p;
          RB.K = BK.p;
          RB.W = BW.p;
          RB.over = over.p;
          launch_fill_u64(bcnt.p, nbt, 0, st());
          launch_fill_i32((i32*)bcur.p, nbt, 0, st());
          launch_fill_i32(over.p, 1, 0, st());
          launch_rl_brow(nlng, bbase.p, brow.p, st());
          launch_rl_fetch(g, tb.p, new_ids.p, RL, RB, st());
          launch_scan_excl((const i64*)bcnt.p, boff.p, nbt, boff.p + nbt, scratch(), st());
          launch_rl_scatter(RL, RB, st());
          launch_rl_dedup(RL, RB, (i32)env_i64("TGP_REP_BUCKET_FILL", REP_BUCKET_FILL), st());
          HIP_TRY(hipMemcpyAsync(C.pinned, over.p, sizeof(i32), hipMemcpyDeviceToHost, st()));
          C.sync();
          bucketed = ((const i32*)C.pinned)[0] == 0;
          BK = DBuf<u64>();
          BW = DBuf<i32>();
          if (bucketed) launch_rl_count(RL, RB, st());
        }
        if (!bucketed) {
          T = DBuf<u64>(&C.pool, ttot);
          W = DBuf<i32>(&C.pool, ttot);
          E = DBuf<i32>(&C.pool, etot);
          launch_fill_u64(T.p, ttot, ~0ull, st());
          launch_fill_i32(W.p, ttot, 0, st());
          RL.toff = toff.p;
          RL.csz = csz.p;
          RL.T = T.p;
          RL.W = W.p;
          RL.E = E.p;
          launch_rep_long(g, tb.p, new_ids.p, RL, 0, nullptr, nullptr, nullptr, st());
        }
        launch_scan_excl(chunk_cnt.p, choff.p, nch, choff.p + nch, scratch(), st());
        launch_rep_rowcnt(RL, cnt.p, st());
      }
      launch_rep_rows(g, tb.p, new_ids.p, n_c, dsz.p, med.p, nmed, meds.p, nmeds, eoff.p, R.p,
                      0, cnt.p, nullptr, nullptr, nullptr, st());
      launch_scan_excl(cnt.p, c->xadj.p, n_c, c->xadj.p + n_c, scratch(), st());
      m_c = read_i64(c->xadj.p + n_c);
      c->adj = DBuf<i32>(&C.pool, std::max<i64>(m_c, 1));
      c->w = DBuf<i32>(&C.pool, std::max<i64>(m_c, 1));
      if (m_c > 0) {
        launch_rep_rows(g, tb.p, new_ids.p, n_c, dsz.p, med.p, nmed, meds.p, nmeds, eoff.p,
                        R.p, 1, nullptr, c->xadj.p, c->adj.p, c->w.p, st());
        if (bucketed) launch_rl_emit(RL, RB, c->xadj.p, c->adj.p, c->w.p, st());
        else launch_rep_long(g, tb.p, new_ids.p, RL, 1, c->xadj.p, c->adj.p, c->w.p, st());
      }
    }
    C.prof.end(KID_CONTRACT, 24.0 * n_c + 12.0 * lv.m + 8.0 * m_c);
    i64 max_edges = 2 * ecount.back();                         // gpu_part.cu:1076-1077
    if (!wide && max_edges > INT_MAX) max_edges = INT_MAX;
    tr.push(TGP_TRACE_LEVEL, (int)L.size() - 1, 0, 0, ncur, lv.m, n_c, m_c, 0);
    if (m_c <= 0 || m_c > max_edges) {                         // gpu_part.cu:1121-1125
      if (S) {
        S->fail_level = (i64)L.size() - 1;
        S->fail_value = m_c;
      }
      set_error("Invalid new_coarsed_edge: " + std::to_string(m_c));
      throw Fail(TGP_EEDGES);
    }
    c->m = m_c;
    // the repaired contraction of a symmetric level is symmetric; the
    // reference's contraction (row r from fine vertex r) is not
    c->symmetric = repair && lv.symmetric;
    // a repaired contraction sums both directions of a weight-symmetric level
    // alike and keeps one entry per neighbour
    c->simple_ws = repair && lv.simple_ws;
    c->tb = std::move(tb);
    c->xadj_p = c->xadj.p;
    c->adj_p = c->adj.p;
    c->w_p = c->w.p;
    L.push_back(std::move(c));
    ecount.push_back(m_c);
    ncur = n_c;
  }
}

LevelAux& Run::aux(Level& lv, i64 nedges) {
  LevelAux& a = lv.aux;
  if (a.valid && a.nedges == nedges) return a;
  a.heavy.ensure(&C.pool, std::max<i64>(lv.n, 1));
  DBuf<i32> cnt32(&C.pool, 2);
  DBuf<u64> maxabs(&C.pool, 1);
  HIP_TRY(hipMemsetAsync(cnt32.p, 0, 2 * sizeof(i32), st()));
  HIP_TRY(hipMemsetAsync(maxabs.p, 0, sizeof(u64), st()));
  C.prof.begin(KID_MISC);
  launch_level_census(lv.dg(), nedges, HEAVY_DEG, a.heavy.p, cnt32.p, cnt32.p + 1, maxabs.p,
                      st());
  C.prof.end(KID_MISC, 12.0 * lv.m);
  HIP_TRY(hipMemcpyAsync(C.pinned, cnt32.p, 2 * sizeof(i32), hipMemcpyDeviceToHost, st()));
  HIP_TRY(hipMemcpyAsync(C.pinned + 1, lv.xadj_p + (lv.n - 1), sizeof(i64),
                         hipMemcpyDeviceToHost, st()));
  C.sync();
  a.nheavy = ((const i32*)C.pinned)[0];
  a.unit = ((const i32*)C.pinned)[1] == 0;
  a.mid.ensure(&C.pool, std::max<i64>(a.nheavy, 1));
  a.medium.ensure(&C.pool, std::max<i64>(a.nheavy, 1));
  a.huge.ensure(&C.pool, std::max<i64>(a.nheavy, 1));
  DBuf<i32> cnt3(&C.pool, 3);
  HIP_TRY(hipMemsetAsync(cnt3.p, 0, 3 * sizeof(i32), st()));
  launch_split_heavy(lv.dg(), nedges, a.heavy.p, a.nheavy, a.mid.p, cnt3.p, a.medium.p,
                     cnt3.p + 1, a.huge.p, cnt3.p + 2, st());
  HIP_TRY(hipMemcpyAsync(C.pinned + 2, cnt3.p, 3 * sizeof(i32), hipMemcpyDeviceToHost, st()));
  C.sync();
  a.nmid = ((const i32*)(C.pinned + 2))[0];
  a.nmedium = ((const i32*)(C.pinned + 2))[1];
  a.nhuge = ((const i32*)(C.pinned + 2))[2];
  a.ctr.ensure(&C.pool, 8 * 32);  // GAIN_CTRS counters, 128 bytes apart
  // integer tallies are exact for every row when no row's sum of |w| passes 2^24
  a.all_exact = 1;
  a.nover = 0;
  if (!a.unit) {
    a.over.ensure(&C.pool, lv.n + 1);
    HIP_TRY(hipMemsetAsync(a.over.p, 0, sizeof(i32), st()));
    launch_row_abs(lv.dg(), nedges, HEAVY_DEG, a.heavy.p, a.nheavy, maxabs.p, a.over.p, st());
    HIP_TRY(hipMemcpyAsync(C.pinned + 3, a.over.p, sizeof(i32), hipMemcpyDeviceToHost, st()));
    C.sync();
    a.nover = ((const i32*)(C.pinned + 3))[0];
    a.all_exact = a.nover == 0;
  }
  // chunks of the long rows, built on the host (nhuge <= m / 4096)
  a.nch = 0;
  if (a.nhuge > 0) {
    DBuf<i64> hb(&C.pool, 2 * a.nhuge);
    launch_huge_bounds(lv.dg(), nedges, a.huge.p, a.nhuge, hb.p, st());
    std::vector<i64> h((size_t)(2 * a.nhuge));
    HIP_TRY(hipMemcpyAsync(h.data(), hb.p, h.size() * sizeof(i64), hipMemcpyDeviceToHost, st()));
    C.sync();
    std::vector<i32> row;
    std::vector<i64> beg, end;
    for (i64 r = 0; r < a.nhuge; r++)
      for (i64 e = h[2 * r]; e < h[2 * r + 1]; e += GAIN_CHUNK_ARCS) {
        row.push_back((i32)r);
        beg.push_back(e);
        end.push_back(std::min(e + GAIN_CHUNK_ARCS, h[2 * r + 1]));
      }
    a.nch = (i64)row.size();
    a.ch_row.ensure(&C.pool, a.nch);
    a.ch_beg.ensure(&C.pool, a.nch);
    a.ch_end.ensure(&C.pool, a.nch);
    HIP_TRY(hipMemcpyAsync(a.ch_row.p, row.data(), a.nch * sizeof(i32), hipMemcpyHostToDevice,
                           st()));
    HIP_TRY(hipMemcpyAsync(a.ch_beg.p, beg.data(), a.nch * sizeof(i64), hipMemcpyHostToDevice,
                           st()));
    HIP_TRY(hipMemcpyAsync(a.ch_end.p, end.data(), a.nch * sizeof(i64), hipMemcpyHostToDevice,
                           st()));
    a.gtally.ensure(&C.pool, a.nhuge * k);
    a.gabs.ensure(&C.pool, a.nhuge);
    HIP_TRY(hipMemsetAsync(a.gtally.p, 0, a.nhuge * k * sizeof(i32), st()));
    HIP_TRY(hipMemsetAsync(a.gabs.p, 0, a.nhuge * sizeof(i64), st()));
    C.sync();  // the host vectors are freed on return
  }
  a.e_end = std::max<i64>(C.pinned[1], nedges);  // rows 0..n-2 whole, last ends at nedges
  a.last_begin = C.pinned[1];
  a.nedges = nedges;
  a.valid = true;
  return a;
}

// the level's arc -> source array, built at its first use (the cut kernel's;
// the passes of a level with kept tallies take their cuts without it)
const i32* Run::arc_source(Level& lv) {
  LevelAux& a = lv.aux;
  if (!a.asrc_ok) {
    a.asrc.ensure(&C.pool, std::max<i64>(lv.m, 1));
    C.prof.begin(KID_MISC);
    launch_arc_source(lv.dg(), a.asrc.p, st());
    C.prof.end(KID_MISC, 4.0 * lv.m + 8.0 * lv.n);
    a.asrc_ok = true;
  }
  return a.asrc.p;
}

i64 Run::compute_cut(Level& lv, i64 nedges, const i32* part) {
  LevelAux& a = aux(lv, nedges);
  C.prof.begin(KID_CUT);
  DBuf<u8> part8;
  if (k <= PART8_MAXK) {  // a byte copy keeps the gathered partition in L2
    part8 = DBuf<u8>(&C.pool, lv.n);
    launch_part8(part, lv.n, k, part8.p, st());
  }
  launch_cut_arcs(lv.dg(), a.e_end, arc_source(lv), part, part8.p, k, a.unit, accs.p, st());
  launch_acc_finish(accs.p, (u64*)(small.p + 1), st());
  C.prof.end(KID_CUT, (a.unit ? 12.0 : 16.0) * a.e_end);
  return wrap_cut((u64)read_i64(small.p + 1));
}

int Run::initial_partition(DBuf<i32>& part, i64* npairs) {
  Level& cl = *L.back();
  const i64 n = cl.n;
  // unique undirected pairs (gpu_part.cu:810-827), counted on the device: on
  // a symmetric level whose rows hold no repeats, half the non-loop arcs;
  // otherwise through a hash table of the pairs
  *npairs = -1;
  if (cl.symmetric && cl.simple_ws) {
    C.prof.begin(KID_PAIRS);
    HIP_TRY(hipMemsetAsync(small.p + 2, 0, sizeof(i64), st()));
    launch_count_pair_arcs(cl.dg(), (u64*)(small.p + 2), st());
    C.prof.end(KID_PAIRS, 4.0 * cl.m + 8.0 * n);
    const i64 arcs = read_i64(small.p + 2);
    if ((arcs & 1) == 0) *npairs = arcs / 2;
  }
  if (*npairs < 0) {
    i64 cap = 1;
    while (cap < 2 * cl.m + 2) cap <<= 1;
    DBuf<u64> table(&C.pool, cap);
    C.prof.begin(KID_PAIRS);
    launch_fill_u64(table.p, cap, ~0ull, st());
    HIP_TRY(hipMemsetAsync(small.p + 2, 0, sizeof(i64), st()));
    launch_pairs(cl.dg(), table.p, cap, (u64*)(small.p + 2), st());
    C.prof.end(KID_PAIRS, 12.0 * cl.m + 8.0 * n);
    *npairs = read_i64(small.p + 2);
  }
  part.ensure(&C.pool, n);
  const auto t0 = std::chrono::steady_clock::now();
  if (O.initpart == tgp_initpart_block) {
    launch_block_init(part.p, n, k, st());
  } else {
    // host seam: the coarsest CSR goes to the callback in host memory
    std::vector<i64> hx((size_t)n + 1);
    std::vector<i32> ha((size_t)std::max<i64>(cl.m, 1)), hw((size_t)std::max<i64>(cl.m, 1));
    HIP_TRY(hipMemcpyAsync(hx.data(), cl.xadj_p, (n + 1) * sizeof(i64), hipMemcpyDeviceToHost,
                           st()));
    if (cl.m > 0) {
      HIP_TRY(hipMemcpyAsync(ha.data(), cl.adj_p, cl.m * sizeof(i32), hipMemcpyDeviceToHost,
                             st()));
      HIP_TRY(hipMemcpyAsync(hw.data(), cl.w_p, cl.m * sizeof(i32), hipMemcpyDeviceToHost, st()));
    }
    C.sync();
    tgp_graph hg{n, cl.m, hx.data(), ha.data(), hw.data()};
    std::vector<i32> hp((size_t)n, 0);
    tgp_initpart_fn fn = O.initpart ? O.initpart : tgp_initpart_gpmetis;
    int rc;
    {
      // one callback at a time: the gpmetis seam uses fixed file names in the
      // working directory (gpu_part.cu:829, 1163), and a caller's callback
      // need not be reentrant; the block partition and the seam in a
      // directory of the item's own run concurrently
      static std::mutex initpart_mu;
      std::unique_lock<std::mutex> lock(initpart_mu, std::defer_lock);
      if (fn != tgp_initpart_block && fn != tgp_initpart_gpmetis_dir) lock.lock();
      rc = fn(&hg, *npairs, k, hp.data(), O.initpart_user);
    }
    if (rc != 0) {
      if (last_error().empty()) set_error("initial partitioner failed");
      throw Fail(rc < 0 ? rc : TGP_EINITPART);
    }
    for (auto& p : hp)
      if (p < 0 || p >= k) p = 0;  // gpu_part.cu:1172
    HIP_TRY(hipMemcpyAsync(part.p, hp.data(), n * sizeof(i32), hipMemcpyHostToDevice, st()));
    C.sync();
  }
  initpart_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

// cut_known = false: *cut is not given; it is the entry partition's cut, which
// the first pass's gain launch sums anyway (the same row ranges as
// compute_cut), read with that pass's status.
void Run::refine(int level, i64 nedges, DBuf<i32>& part, DBuf<i32>& pc, i64* cut,
                 bool cut_known) {
  const int TOKEN_LANES = token_lanes(), GAIN_LANES = gain_lanes();
  Level& lv = *L[level];
  const i64 n = lv.n;
  const DGraph g = lv.dg();
  const i64 T = n / k;
  const float Tf = (float)T;
  const float maxImbalance = 0.05f;

  LevelAux& ax = aux(lv, nedges);
  const i32* heavy = ax.heavy.p;
  const i64 nheavy = ax.nheavy;
  const int unit = ax.unit;

  if (!cut_known && n / k == 0) {
    *cut = compute_cut(lv, nedges, part.p);
    cut_known = true;
  }
  bool need_init = !cut_known;  // the first pass also reads the entry cut
  i64 bestCut = *cut;
  float bestScore = 1e9f;
  DBuf<i32> best(&C.pool, n);
  HIP_TRY(hipMemcpyAsync(best.p, part.p, n * sizeof(i32), hipMemcpyDeviceToDevice, st()));
  int passes = 0;

  if (T == 0) {
    // Every gain is -inf or NaN (division by T = 0), so no vertex becomes a
    // candidate and both phases stop at their first pass with no moves.
    passes = 2;
  } else {
    DBuf<float> gain(&C.pool, n), mv_gain(&C.pool, n), sk_gain(&C.pool, n);
    DBuf<i32> target(&C.pool, n), mv_v(&C.pool, n), mv_src(&C.pool, n), mv_dst(&C.pool, n),
        sk_idx(&C.pool, n), sdst(&C.pool, n);
    DBuf<float> cand(&C.pool, n);
    // byte copy of part for the gathers of the gain kernels (1 byte per vertex
    // keeps the partition resident in L2); rebuilt after every bulk change of part
    DBuf<u8> part8;
    if (k <= PART8_MAXK) {
      part8 = DBuf<u8>(&C.pool, n);
      launch_part8(part.p, n, k, part8.p, st());
    }
    // Kept tallies (n x k ints) on simple weight-symmetric levels with exact
    // integer tallies: built by the first gain launch after any bulk change of
    // part, then updated from the applied moves of each commit.  The table
    // may take up to 40% of the free device memory.
    size_t free_b = 0, total_b = 0;
    if (C.pool.arena()) {
      free_b = C.pool.arena_free();
    } else {
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        free_b = 0;
      }
      free_b += C.pool.cached_bytes();
    }
    // A few rows whose sum of |w| passes 2^24 keep the kept-tally form: their
    // gains are gathered again in fp32 row order after every tally pass.
    // Rows whose weights sum to at most 255 keep bytes, the others ints
    // (TallyRef); a level where over half the rows need ints keeps ints for
    // every row (as does TGP_TALLY_BYTES=0).
    bool tmode = k <= 512 && (ax.all_exact || ax.nover <= OVER_ROWS_MAX) && lv.simple_ws &&
                 env_i64("TGP_KEEP_TALLIES", 1) != 0;
    DBuf<i32> T32, bix, bigv, app, napp;
    DBuf<u32> bigmask;
    DBuf<u8> T8;
    DBuf<u16> T16;
    TallyRef TR;
    i64 nbig = n;
    bool all_int = true;
    // byte rows pay off with one lane per vertex (k <= 64, gain lanes 1),
    // short rows with 4 lanes per vertex for 64 < k <= TALLY16_MAXK; the
    // kernels for other widths read int rows only
    const bool bytes = env_i64("TGP_TALLY_BYTES", 1) != 0 && k <= 64 && GAIN_LANES == 1;
    const bool shorts =
        env_i64("TGP_TALLY_SHORTS", 1) != 0 && k > 64 && k <= TALLY16_MAXK && GAIN_LANES == 1;
    if (tmode && (bytes || shorts)) {
      bix = DBuf<i32>(&C.pool, n);
      bigmask = DBuf<u32>(&C.pool, (n + 31) / 32 + 2);
      DBuf<i32> nb(&C.pool, 1);
      HIP_TRY(hipMemsetAsync(nb.p, 0, sizeof(i32), st()));
      launch_tally_class(g, nedges, unit, 0, shorts ? 65535 : 255, bix.p, bigmask.p, nb.p, st());
      HIP_TRY(hipMemcpyAsync(C.pinned, nb.p, sizeof(i32), hipMemcpyDeviceToHost, st()));
      C.sync();
      nbig = ((const i32*)C.pinned)[0];
      all_int = 2 * nbig > n;
      if (all_int) {
        nbig = n;
        bix = DBuf<i32>();
        bigmask = DBuf<u32>();
      }
    }
    // bytes one pass reads from the kept tallies (and the row classes)
    const double tally_bytes =
        all_int ? (double)n * tally_row_stride(k) * 4
                : (double)n * (shorts ? 2 * tally16_stride(k) : tally8_stride(k)) +
                      (double)nbig * tally_row_stride(k) * 4 + n / 8.0;
    if (tmode) tmode = tally_bytes <= 0.4 * (double)free_b;
    bool T_valid = false;
    const bool last_full = nedges == lv.m;
    const bool last_partial = !last_full && nedges > ax.last_begin;
    if (tmode) {
      // row padding (entries past k) stays zero: the updates touch parts < k
      if (!all_int && shorts) {
        T16 = DBuf<u16>(&C.pool, n * tally16_stride(k));
        if (k % 8) HIP_TRY(hipMemsetAsync(T16.p, 0, (size_t)n * tally16_stride(k) * 2, st()));
      } else if (!all_int) {
        T8 = DBuf<u8>(&C.pool, n * tally8_stride(k));
        if (k % 16) HIP_TRY(hipMemsetAsync(T8.p, 0, (size_t)n * tally8_stride(k), st()));
      }
      if (!all_int) {
        bigv = DBuf<i32>(&C.pool, std::max<i64>(nbig, 1));
        launch_big_list(bix.p, n, bigv.p, st());
      }
      T32 = DBuf<i32>(&C.pool, std::max<i64>(nbig, 1) * tally_row_stride(k));
      if (k % 4)
        HIP_TRY(hipMemsetAsync(T32.p, 0, (size_t)std::max<i64>(nbig, 1) * tally_row_stride(k) * 4,
                               st()));
      TR = TallyRef{T8.p, T32.p, bix.p, bigmask.p, all_int ? 1 : 0, bigv.p, (i32)nbig, T16.p};
      app = DBuf<i32>(&C.pool, n);
      napp = DBuf<i32>(&C.pool, 1);
    }
    DBuf<u8> sel(&C.pool, n), src_big(&C.pool, k);
    DBuf<u8> flag(&C.pool, n);
    // token witnesses of the level's passes (-1: none yet)
    DBuf<i32> wit;
    if (env_i64("TGP_TOKEN_WITNESS", 1) != 0) {
      wit = DBuf<i32>(&C.pool, n);
      HIP_TRY(hipMemsetAsync(wit.p, 0xFF, (size_t)n * sizeof(i32), st()));
    }
    DBuf<i32> bcnt(&C.pool, compaction_blocks()), boff(&C.pool, compaction_blocks());
    DBuf<i64> avail(&C.pool, k);
    DBuf<float> bg(&C.pool, 2 * (size_t)k);
    // outc[0..k) moves out of each part, outc[k..2k) moves into each part;
    // fit[0] = 1 while every destination's moves fit its capacity
    DBuf<i32> outc(&C.pool, 2 * (size_t)k), ctr(&C.pool, 4), fit(&C.pool, 1);
    DBuf<i32> H(&C.pool, select_scratch_elems(n, k));
    // the move sort is skipped when it cannot change the selection: all M
    // moves fit their destinations and the network would sort them fully
    const bool sort_skip = env_i64("TGP_SORT_SKIP", 1) != 0;
    // TGP_SORT_AHEAD=0: the sort's first pass waits for the host's decision
    const bool sort_ahead = env_i64("TGP_SORT_AHEAD", 1) != 0;
    bool last_skipped = false;
    DBuf<i32> gscr(&C.pool, std::max<i64>(gain_gscratch_elems(k), 1));
    std::vector<i32> hpc((size_t)k);
    // a pass's count, cut, selection counters and part counts go to the host
    // in one copy when they fit the pinned staging area
    const bool pack = 6 + (i64)k <= 2 * (i64)Context::PINNED_ELEMS;
    DBuf<i64> pstat(&C.pool, pack ? 3 + ((size_t)k + 1) / 2 : 1);
    const double arc_b = unit ? 8.0 : 12.0;

    DBuf<u64> cutacc(&C.pool, 1);
    i64 lastM = n;  // moves buffered by the last pass of this level (token lane choice)
    bool best_cur = true;  // part equals best (no commit since best was recorded)
    for (int phase = 0; phase < 2; phase++) {
      const int maxPass = phase ? 40 : 30;               // gpu_part.cu:545, 647
      bool balanced = false;
      if (O.pass_dump) dump_part(level, phase, -1, part.p, n);
      // A pass's edge cut is the cut of the partition the next pass's gain
      // kernel reads, so that kernel tallies it and each pass costs one host
      // sync.  The next pass is therefore started before the previous pass's
      // stop test (gpu_part.cu:632-640, 736-744) is known; until its commit it
      // changes nothing, and it is dropped when that test ends the phase.
      bool pending = false;  // the previous pass has committed; its cut, S and pc are in flight
      i64 prevM = 0;
      int prevPass = -1;
      for (int pass = 0;; pass++) {
        const bool run = pass < maxPass;
        i32* ctr_cur = ctr.p + 2 * (pass & 1);
        i32* ctr_prev = ctr.p + 2 * ((pass + 1) & 1);
        if (run) {
          C.prof.begin(KID_MISC);
          // also clears outc[0..2k) and the pass's two selection counters, sets fit
          launch_bg_table(pc.p, k, n, phase, bg.p, src_big.p, outc.p, 2 * k, ctr_cur, fit.p,
                          st());
          C.prof.end(KID_MISC, 0);
          C.prof.begin(KID_GAIN);
          bool fused;
          if (tmode && T_valid) {
            if (last_partial) launch_tally_last(g, nedges, part.p, TR, k, unit, st());
            launch_gain_tallies(n, part.p, TR, k, bg.p, src_big.p, gain.p, target.p, cand.p,
                                phase, (pending || need_init) ? accs.p : nullptr, GAIN_LANES,
                                st());
            launch_gain_rows(g, nedges, part.p, k, bg.p, src_big.p, gain.p, target.p, cand.p,
                             phase, unit, ax.over.p + 1, ax.nover, st());
            fused = true;
            C.prof.end(KID_GAIN, tally_bytes + 17.0 * n);
          } else {
            fused = launch_gain(g, nedges, part.p, part8.p, k, bg.p, src_big.p, gain.p, target.p,
                                cand.p, phase, unit, gscr.p, ax.lists(),
                                (pending || need_init) ? accs.p : nullptr, tmode ? &TR : nullptr,
                                st());
            T_valid = tmode;
            C.prof.end(KID_GAIN, arc_b * lv.m + 21.0 * n);
          }
          if ((pending || need_init) && !fused) {
            C.prof.begin(KID_CUT);
            launch_cut_arcs(g, ax.e_end, arc_source(lv), part.p, part8.p, k, unit, accs.p, st());
            C.prof.end(KID_CUT, (unit ? 12.0 : 16.0) * ax.e_end);
          }
          if (pending || need_init) launch_acc_finish(accs.p, cutacc.p, st());
          C.prof.begin(KID_TOKEN);
          // Light rows: the arc stream (lanes 0) pays off when most candidates
          // win and read their whole rows; with few winners (under 30 % of the
          // vertices moved in the last pass) one thread per row, which stops at
          // the first losing neighbour, measured faster.  Both are exact.
          const int tlanes = (TOKEN_LANES == 0 && 10 * lastM < 3 * n) ? 1 : TOKEN_LANES;
          launch_token(g, nedges, cand.p, part.p, target.p, k, flag.p, heavy, nheavy, HEAVY_DEG,
                       tlanes, wit.p, st());
          // 9m + (o + 6)n: arcs and the candidates' gains, offsets, own gain and flag
          C.prof.end(KID_TOKEN, 9.0 * lv.m + 14.0 * n);
          C.prof.begin(KID_COMPACT);
          launch_compact_moves(flag.p, n, part.p, target.p, gain.p, k, bcnt.p, boff.p, small.p,
                               mv_v.p, mv_src.p, mv_dst.p, mv_gain.p, outc.p, outc.p + k, st());
          C.prof.end(KID_COMPACT, 2.0 * n);
          // capacities from this pass's counts (the previous commit is done)
          C.prof.begin(KID_MISC);
          launch_capacity(pc.p, outc.p, k, n, phase, avail.p, tmode ? napp.p : nullptr,
                          outc.p + k, fit.p, st());
          C.prof.end(KID_MISC, 0);
          if (!pack)
            HIP_TRY(hipMemcpyAsync(C.pinned, small.p, sizeof(i64), hipMemcpyDeviceToHost, st()));
        } else if (pending) {
          C.prof.begin(KID_CUT);
          if (tmode && T_valid) {
            // the tallies are current after the last commit (the last row's
            // partial range is rebuilt as the gain pass does)
            if (last_partial) launch_tally_last(g, nedges, part.p, TR, k, unit, st());
            launch_tally_cut(n, part.p, TR, k, accs.p, st());
            launch_acc_finish(accs.p, cutacc.p, st());
            C.prof.end(KID_CUT, tally_bytes + 4.0 * n);
          } else {
            launch_cut_arcs(g, ax.e_end, arc_source(lv), part.p, part8.p, k, unit, accs.p, st());
            launch_acc_finish(accs.p, cutacc.p, st());
            C.prof.end(KID_CUT, (unit ? 12.0 : 16.0) * ax.e_end);
          }
        } else {
          break;
        }
        if (pack) {
          launch_pass_status(run ? (const i64*)small.p : nullptr,
                             (pending || need_init) ? cutacc.p : nullptr,
                             pending ? ctr_prev : nullptr, pending ? pc.p : nullptr, k,
                             run ? fit.p : nullptr, pstat.p, st());
          const size_t bytes = pending ? 24 + 4 * (size_t)k : need_init ? 16 : 8;
          HIP_TRY(hipMemcpyAsync(C.pinned, pstat.p, bytes, hipMemcpyDeviceToHost, st()));
        } else if (pending || need_init) {
          HIP_TRY(hipMemcpyAsync(C.pinned + 1, cutacc.p, sizeof(u64), hipMemcpyDeviceToHost,
                                 st()));
        }
        if (!pack && pending) {
          HIP_TRY(hipMemcpyAsync(C.pinned + 2, ctr_prev, 2 * sizeof(i32), hipMemcpyDeviceToHost,
                                 st()));
          HIP_TRY(hipMemcpyAsync(hpc.data(), pc.p, k * sizeof(i32), hipMemcpyDeviceToHost, st()));
        }
        // The first tile pass of this pass's move sort (M read on the device)
        // runs while the host waits for the status and does the bookkeeping;
        // it writes only the sort's buffers, which nothing reads when the host
        // then skips the sort or ends the phase.  Not after a pass whose sort
        // was skipped (the next one likely skips too).
        const bool ahead = run && sort_ahead && !last_skipped;
        if (ahead) {
          C.mark();
          C.prof.begin(KID_SORT);
          launch_sort_first_dev(mv_gain.p, small.p, n, sk_gain.p, sk_idx.p, st());
          C.prof.end(KID_SORT, 0);
          C.wait_mark();
        } else {
          C.sync();
        }
        if (pack && pending) std::memcpy(hpc.data(), (const i32*)C.pinned + 6, k * sizeof(i32));
        if (need_init) {  // the entry partition's cut, summed by this pass's gain launch
          need_init = false;
          bestCut = wrap_cut((u64)C.pinned[1]);
          *cut = bestCut;
        }
        if (pending) {
          // the previous pass's bookkeeping, gpu_part.cu:605-640 / 707-744
          pending = false;
          const i64 cur = wrap_cut((u64)C.pinned[1]);
          const i32 Ssel = ((const i32*)(C.pinned + 2))[0];
          const i32 pads = ((const i32*)(C.pinned + 2))[1];
          float score = 0.0f;                              // gpu_part.cu:621-626, 723-731
          balanced = true;
          for (i32 q = 0; q < k; q++) {
            const float diff = (float)hpc[q] - Tf;
            const float sq = diff * diff;
            score = score + sq;
            if (std::fabs(diff) > Tf * maxImbalance) balanced = false;
          }
          score = std::sqrt(score / (float)k);
          if (cur < bestCut || ((float)cur <= (float)bestCut * 1.01f && score < bestScore)) {
            bestCut = cur;
            bestScore = score;
            HIP_TRY(hipMemcpyAsync(best.p, part.p, n * sizeof(i32), hipMemcpyDeviceToDevice,
                                   st()));
            best_cur = true;
          }
          *cut = cur;
          tr.push(TGP_TRACE_PASS, level, phase, prevPass, prevM, Ssel, cur, pads,
                  O.check_level ? hash_dev(part.p, n) : 0);
          if (O.pass_dump) dump_part(level, phase, prevPass, part.p, n);
          const bool quiet = (prevM < n / 200) && (Ssel < n / 400);  // gpu_part.cu:634, 738
          if (quiet && (phase == 0 || balanced)) break;
          if (phase && balanced) break;                    // loop test, gpu_part.cu:647
        }
        if (!run) break;
        const i64 M = pack ? std::min<i64>(((const i32*)C.pinned)[0], n)
                           : std::min<i64>(C.pinned[0], n);  // gpu_part.cu:585
        const bool all_fit = pack && ((const i32*)C.pinned)[1] != 0;
        passes++;
        if (M <= 0) break;                                 // gpu_part.cu:638-640, 742-744
        i64 len = M;                                       // gpu_part.cu:487-497
        i64 padded = 1;
        while (padded < len) padded <<= 1;
        const bool full_network = padded <= n;
        if (padded > n) padded = n;
        if (padded > len) len = padded;
        if (sort_skip && all_fit && full_network) {
          // every move is selected whatever the network's order, and with a
          // full network no pad lands among the first M (gains of buffered
          // moves exceed the pads' -1e9)
          C.prof.begin(KID_SELECT);
          launch_select_all(M, sk_idx.p, sel.p, ctr_cur, ctr_cur + 1, st());
          C.prof.end(KID_SELECT, 5.0 * M);
          sorts_skipped++;
          last_skipped = true;
        } else {
          last_skipped = false;
          C.prof.begin(KID_SORT);
          launch_sort_moves(mv_gain.p, M, len, sk_gain.p, sk_idx.p, st(), ahead);
          C.prof.end(KID_SORT, 16.0 * (double)len);
          C.prof.begin(KID_SELECT);
          launch_select(sk_idx.p, mv_dst.p, M, k, avail.p, sdst.p, H.p, sel.p, ctr_cur,
                        ctr_cur + 1, st());
          C.prof.end(KID_SELECT, 13.0 * M);
        }
        C.prof.begin(KID_COMMIT);
        launch_commit(sk_idx.p, sel.p, M, mv_v.p, mv_src.p, mv_dst.p, part.p, part8.p, pc.p, k,
                      tmode ? app.p : nullptr, napp.p, st());
        if (tmode && T_valid)
          launch_tally_update(g, TR, k, app.p, napp.p, M, mv_v.p, mv_src.p, mv_dst.p, unit,
                              last_full ? 0 : 1, st());
        C.prof.end(KID_COMMIT, 17.0 * M);
        best_cur = false;
        pending = true;
        prevM = M;
        lastM = M;
        prevPass = pass;
      }
      // restore the best partition (gpu_part.cu:642, 746); when nothing
      // was committed since it was recorded, part already equals it and the
      // kept tallies and the byte copy stay valid
      if (!best_cur) {
        HIP_TRY(hipMemcpyAsync(part.p, best.p, n * sizeof(i32), hipMemcpyDeviceToDevice, st()));
        if (part8.p) launch_part8(part.p, n, k, part8.p, st());
        T_valid = false;
      }
      if (phase == 0) *cut = bestCut;
    }
  }
  // gpu_part.cu:749 recomputes the cut of the restored best partition: that
  // partition's cut is bestCut (the cut taken when it was recorded, or the
  // entry cut, with the same range rules)
  *cut = bestCut;
  passes_total += passes;
  tr.push(TGP_TRACE_REFINE, level, 0, passes, n, nedges, *cut, 0,
          O.check_level ? hash_dev(part.p, n) : 0);
}

// Learned passes at one level (tgp_learned_model): GNN scores -> each
// vertex's best other part, admitted only with a positive exact integer gain;
// then the reference's token rule on those gains, move buffer, capacity,
// sort, selection and commit (gpu_part.cu:316-397, 472-506).  Stops when
// fewer than min_admit * n moves are admitted.
void Run::learned_passes(int level, i64 nedges, DBuf<i32>& part, DBuf<i32>& pc, i64 cap) {
  const int TOKEN_LANES = token_lanes();
  const tgp_learned_model& Mdl = *O.learned;
  Level& lv = *L[level];
  const i64 n = lv.n;
  if (Mdl.k != k || k > 256 || Mdl.hidden <= 0 || Mdl.features < 2 * k + 2) {
    set_error("learned model does not fit: model k, k <= 256, hidden > 0, features >= 2k+2");
    throw Fail(TGP_EINVAL);
  }
  if (n < 2) return;
  LevelAux& ax = aux(lv, nedges);
  const DGraph g = lv.dg();
  const int F = Mdl.features, Hd = Mdl.hidden;
  typedef uint16_t u16;
  DBuf<u16> X(&C.pool, n * F), Pb(&C.pool, n * Hd), Hb(&C.pool, n * Hd), M1(&C.pool, n * Hd);
  DBuf<float> Pf(&C.pool, n * Hd), Hf(&C.pool, n * Hd), Z(&C.pool, n * k);
  DBuf<i32> tally(&C.pool, n * k);
  DBuf<float> gain(&C.pool, n), mv_gain(&C.pool, n), sk_gain(&C.pool, n);
  DBuf<i32> target(&C.pool, n), mv_v(&C.pool, n), mv_src(&C.pool, n), mv_dst(&C.pool, n),
      sk_idx(&C.pool, n), sdst(&C.pool, n), outc(&C.pool, k), ctr(&C.pool, 2);
  DBuf<i32> H(&C.pool, select_scratch_elems(n, k));
  DBuf<u8> sel(&C.pool, n);
  DBuf<u8> flag(&C.pool, n);
  DBuf<i32> bcnt(&C.pool, compaction_blocks()), boff(&C.pool, compaction_blocks());
  DBuf<i64> avail(&C.pool, k);
  void* s = (void*)st();
  auto ck = [&](int rc) {
    if (rc != 0) {
      set_error("learned pass: kernel launch failed");
      throw Fail(TGP_EHIP);
    }
  };
  const i64 min_moves = std::max<i64>(1, (i64)(Mdl.min_admit * (double)n));
  const int every = std::max(1, Mdl.infer_every);
  for (int p = 0; p < Mdl.max_passes; p++) {
    ck(tgpl_features(n, g.xadj, g.adj, g.w, part.p, k, F, X.p, tally.p, s));
    if (p % every == 0) {
      // H = ReLU(X W1a + mean_N(X W1b) + b1): the mean over the 64 columns of X W1b
      ck(tgpl_gemm(X.p, F, 0, Mdl.W1b, Hd, 0, Pf.p, Hd, Pb.p, (int)n, Hd, F, 1.0f, 0.0f, nullptr,
                   0, s));
      ck(tgpl_mean_agg(n, g.xadj, g.adj, Pb.p, Hd, 0, M1.p, Hf.p, s));
      ck(tgpl_gemm(X.p, F, 0, Mdl.W1a, Hd, 0, Hf.p, Hd, Hb.p, (int)n, Hd, F, 1.0f, 1.0f, Mdl.b1,
                   1, s));
      ck(tgpl_mean_agg(n, g.xadj, g.adj, Hb.p, Hd, 0, M1.p, nullptr, s));
      ck(tgpl_gemm(Hb.p, Hd, 0, Mdl.W2a, k, 0, Z.p, k, nullptr, (int)n, k, Hd, 1.0f, 0.0f, nullptr,
                   0, s));
      ck(tgpl_gemm(M1.p, Hd, 0, Mdl.W2b, k, 0, Z.p, k, nullptr, (int)n, k, Hd, 1.0f, 1.0f, Mdl.b2,
                   0, s));
    }
    ck(tgpl_propose(n, k, Z.p, part.p, tally.p, Mdl.margin, gain.p, target.p, s));
    HIP_TRY(hipMemsetAsync(outc.p, 0, k * sizeof(i32), st()));
    HIP_TRY(hipMemsetAsync(ctr.p, 0, 2 * sizeof(i32), st()));
    launch_token(g, nedges, gain.p, part.p, target.p, k, flag.p, ax.heavy.p, ax.nheavy,
                 HEAVY_DEG, TOKEN_LANES, nullptr, st());
    launch_compact_moves(flag.p, n, part.p, target.p, gain.p, k, bcnt.p, boff.p, small.p,
                         mv_v.p, mv_src.p, mv_dst.p, mv_gain.p, outc.p, nullptr, st());
    HIP_TRY(hipMemcpyAsync(C.pinned, small.p, sizeof(i64), hipMemcpyDeviceToHost, st()));
    C.sync();
    const i64 M = std::min<i64>(C.pinned[0], n);
    if (M < min_moves) break;
    if (cap > 0) launch_bal_avail(pc.p, k, cap, avail.p, st());
    else launch_capacity(pc.p, outc.p, k, n, 0, avail.p, nullptr, nullptr, nullptr, st());
    i64 len = M, padded = 1;
    while (padded < len) padded <<= 1;
    if (padded > n) padded = n;
    if (padded > len) len = padded;
    launch_sort_moves(mv_gain.p, M, len, sk_gain.p, sk_idx.p, st());
    launch_select(sk_idx.p, mv_dst.p, M, k, avail.p, sdst.p, H.p, sel.p, ctr.p, ctr.p + 1, st());
    launch_commit(sk_idx.p, sel.p, M, mv_v.p, mv_src.p, mv_dst.p, part.p, nullptr, pc.p, k,
                  nullptr, nullptr, st());
    passes_total++;
  }
}

// The balance stage of the finest level (tgp_options.balance_eps): balance
// rounds until every part holds at most cap, the learned passes (when a
// model is given) within cap, then refine passes of positive exact gains
// within cap.  No counterpart in the reference.
void Run::balance_finest(DBuf<i32>& part, DBuf<i32>& pc) {
  static const int REFINE_PASSES = 30;
  static const int MAX_ROUNDS = 1000;
  Level& lv = *L[0];
  const i64 n = lv.n;
  const DGraph g = lv.dg();
  const i64 cap = (i64)std::floor((1.0 + (double)O.balance_eps) * (double)((n + k - 1) / k));
  DBuf<i32> T(&C.pool, (size_t)n * k), target(&C.pool, n), gain(&C.pool, n), room(&C.pool, k),
      lim(&C.pool, k), bcnt(&C.pool, 4096), vals(&C.pool, n), vals2(&C.pool, n);
  DBuf<u8> cand(&C.pool, n), keep(&C.pool, n), keep2(&C.pool, n);
  DBuf<u64> keys(&C.pool, n), keys2(&C.pool, n), moved(&C.pool, 1);
  DBuf<i64> total(&C.pool, 1), start(&C.pool, k);
  const i64 sb = std::max<i64>(balance_sort_bytes(n), 16);
  DBuf<u8> sort_tmp(&C.pool, sb);
  const BalScratch S{bcnt.p, total.p, keys.p, keys2.p, vals.p, vals2.p, start.p, sort_tmp.p, sb};
  // the counts of the partition itself (the reference's counts are those of
  // its last pass, not of the best partition it restores)
  HIP_TRY(hipMemsetAsync(pc.p, 0, k * sizeof(i32), st()));
  launch_sanitize_hist(part.p, n, k, pc.p, st());
  std::vector<i32> hpc((size_t)k), hroom((size_t)k), hlim((size_t)k);
  auto counts = [&]() {
    HIP_TRY(hipMemcpyAsync(hpc.data(), pc.p, k * sizeof(i32), hipMemcpyDeviceToHost, st()));
    C.sync();
    bool over = false;
    for (i32 q = 0; q < k; q++) {
      hroom[q] = (i32)std::max<i64>(std::min<i64>(cap - hpc[q], INT_MAX), INT_MIN);
      hlim[q] = (i32)std::max<i64>(hpc[q] - cap, 0);
      over |= hpc[q] > cap;
    }
    HIP_TRY(hipMemcpyAsync(room.p, hroom.data(), k * sizeof(i32), hipMemcpyHostToDevice, st()));
    HIP_TRY(hipMemcpyAsync(lim.p, hlim.data(), k * sizeof(i32), hipMemcpyHostToDevice, st()));
    return over;
  };
  // tallies kept up to date by each round's moves on levels whose reverse
  // arcs carry the same weight; rebuilt every round elsewhere
  const bool incremental = lv.simple_ws;
  bool T_valid = false;
  auto tally = [&]() {
    if (!(incremental && T_valid)) launch_bal_tally(g, part.p, k, T.p, st());
    T_valid = true;
  };
  auto apply = [&](const u8* kp) {
    HIP_TRY(hipMemsetAsync(moved.p, 0, sizeof(u64), st()));
    if (incremental) launch_bal_tally_update(g, kp, part.p, target.p, k, T.p, st());
    launch_bal_apply(kp, n, target.p, part.p, pc.p, moved.p, st());
    HIP_TRY(hipMemcpyAsync(C.pinned, moved.p, sizeof(u64), hipMemcpyDeviceToHost, st()));
    C.sync();
    return (i64)C.pinned[0];
  };
  C.prof.begin(KID_MISC);
  i64 bal_moves = 0, ref_moves = 0;
  int rounds = 0;
  while (rounds < MAX_ROUNDS && counts()) {
    tally();
    launch_bal_best(n, part.p, T.p, k, room.p, 0, cand.p, target.p, gain.p, st());
    launch_bal_rank_keep(cand.p, n, target.p, gain.p, k, room.p, keep.p, S, st());
    launch_bal_rank_keep(keep.p, n, part.p, gain.p, k, lim.p, keep2.p, S, st());
    const i64 mv = apply(keep2.p);
    rounds++;
    bal_moves += mv;
    if (mv == 0) break;
  }
  C.prof.end(KID_MISC, 0);
  if (O.learned) {
    learned_passes(0, lv.m, part, pc, cap);
    T_valid = false;
  }
  C.prof.begin(KID_MISC);
  for (int pass = 0; pass < REFINE_PASSES; pass++) {
    counts();
    tally();
    launch_bal_best(n, part.p, T.p, k, room.p, 1, cand.p, target.p, gain.p, st());
    launch_bal_token(g, cand.p, gain.p, keep.p, st());
    launch_bal_rank_keep(keep.p, n, target.p, gain.p, k, room.p, keep2.p, S, st());
    const i64 mv = apply(keep2.p);
    ref_moves += mv;
    if (mv == 0) break;
  }
  C.prof.end(KID_MISC, 0);
  counts();
  i32 largest = 0;
  for (i32 q = 0; q < k; q++) largest = std::max(largest, hpc[q]);
  const i64 cut = compute_cut(lv, lv.m, part.p);
  tr.push(TGP_TRACE_BALANCE, 0, 0, rounds, bal_moves, ref_moves, cut, largest,
          O.check_level ? hash_dev(part.p, n) : 0);
}

int Run::balance_only(const tgp_graph* g, i32* part_io, i32* pc_out, i64* cut_out) {
  const auto t_start = std::chrono::steady_clock::now();
  const i64 n = g->n, m = g->m;
  setup_level0(g);
  check_symmetric(*L[0]);
  DBuf<i32> part(&C.pool, n), pc(&C.pool, k);
  HIP_TRY(hipMemcpyAsync(part.p, part_io, n * sizeof(i32), hipMemcpyHostToDevice, st()));
  balance_finest(part, pc);
  const i64 cut = compute_cut(*L[0], m, part.p);
  HIP_TRY(hipMemcpyAsync(part_io, part.p, n * sizeof(i32), hipMemcpyDeviceToHost, st()));
  HIP_TRY(hipMemcpyAsync(pc_out, pc.p, k * sizeof(i32), hipMemcpyDeviceToHost, st()));
  C.sync();
  *cut_out = cut;
  if (S) {
    S->seconds_total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    S->trace_len = tr.len;
  }
  return TGP_OK;
}

void Run::setup_level0(const tgp_graph* g) {
  const i64 n = g->n, m = g->m;
  scan_sc = DBuf<i64>(&C.pool, scan_scratch_elems(std::max<i64>(n, 2) * 2 + 64));
  small = DBuf<i64>(&C.pool, 16);
  accs = DBuf<u64>(&C.pool, ACC_ELEMS);
  HIP_TRY(hipMemsetAsync(accs.p, 0, ACC_ELEMS * sizeof(u64), st()));

  // level 0
  std::unique_ptr<Level> l0(new Level());
  l0->n = n;
  l0->m = m;
  if (O.device_input) {
    // a device CSR is checked on the device before any kernel follows its
    // offsets (a host CSR was checked by validate_host_graph)
    DBuf<i32> bad(&C.pool, 1);
    HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(i32), st()));
    launch_csr_check(g->xadj, g->adjncy, n, m, bad.p, st());
    HIP_TRY(hipMemcpyAsync(C.pinned, bad.p, sizeof(i32), hipMemcpyDeviceToHost, st()));
    C.sync();
    if (((const i32*)C.pinned)[0] != 0) {
      set_error("malformed device CSR: need xadj[0]=0, non-decreasing offsets, xadj[n]=m and "
                "every adjncy entry in [0, n)");
      throw Fail(TGP_EINVAL);
    }
    l0->xadj_p = g->xadj;
    l0->adj_p = g->adjncy;
    l0->w_p = g->adjwgt;
  } else {
    l0->xadj = DBuf<i64>(&C.pool, n + 1);
    l0->adj = DBuf<i32>(&C.pool, m);
    l0->w = DBuf<i32>(&C.pool, m);
    HIP_TRY(hipMemcpyAsync(l0->xadj.p, g->xadj, (n + 1) * sizeof(i64), hipMemcpyHostToDevice,
                           st()));
    HIP_TRY(hipMemcpyAsync(l0->adj.p, g->adjncy, m * sizeof(i32), hipMemcpyHostToDevice, st()));
    HIP_TRY(hipMemcpyAsync(l0->w.p, g->adjwgt, m * sizeof(i32), hipMemcpyHostToDevice, st()));
    l0->xadj_p = l0->xadj.p;
    l0->adj_p = l0->adj.p;
    l0->w_p = l0->w.p;
  }
  L.push_back(std::move(l0));
  ecount.push_back(m);
}

int Run::go(const tgp_graph* g, i32* part_out, i32* pc_out, i64* cut_out) {
  const auto t_start = std::chrono::steady_clock::now();
  const i64 n = g->n, m = g->m;
  setup_level0(g);
  coarsen();
  const int pass = (int)L.size() - 1;

  DBuf<i32> part, pc(&C.pool, k);
  i64 npairs = 0;
  initial_partition(part, &npairs);
  HIP_TRY(hipMemsetAsync(pc.p, 0, k * sizeof(i32), st()));
  C.prof.begin(KID_PROJECT);
  launch_sanitize_hist(part.p, L[pass]->n, k, pc.p, st());
  C.prof.end(KID_PROJECT, 8.0 * L[pass]->n);
  i64 cut = compute_cut(*L[pass], npairs, part.p);  // gpu_part.cu:1191
  tr.push(TGP_TRACE_INIT, pass, 0, 0, L[pass]->n, npairs, cut, 0,
          O.check_level ? hash_dev(part.p, L[pass]->n) : 0);
  refine(pass, npairs, part, pc, &cut);                   // gpu_part.cu:1200 (B5)

  for (int level = pass - 1; level >= 0; level--) {       // gpu_part.cu:1205-1274
    const i64 nf = L[level]->n, ncs = L[level + 1]->n;
    DBuf<i32> pf(&C.pool, nf);
    C.prof.begin(KID_PROJECT);
    launch_fill_i32(pf.p, nf, -1, st());
    launch_project(L[level + 1]->tb.p, part.p, ncs, pf.p, nf, k, st());
    HIP_TRY(hipMemsetAsync(pc.p, 0, k * sizeof(i32), st()));
    launch_sanitize_hist(pf.p, nf, k, pc.p, st());
    C.prof.end(KID_PROJECT, 12.0 * ncs + 12.0 * nf);
    part = std::move(pf);
    if (O.learned && level == 0 && !(O.balance_eps > 0.0f))
      learned_passes(level, ecount[level], part, pc);
    // the entry cut of the level comes with its first pass (the learned passes
    // before level 0's refinement leave it to compute_cut)
    const bool learned0 = O.learned && level == 0 && !(O.balance_eps > 0.0f);
    if (learned0) cut = compute_cut(*L[level], ecount[level], part.p);
    refine(level, ecount[level], part, pc, &cut, learned0);
  }
  if (O.balance_eps > 0.0f) balance_finest(part, pc);
  // gpu_part.cu:1285: the level-0 refinement ended with this partition's cut
  // over the same ranges unless level 0 is the coarsest level (B5 ranges there)
  // or the balance stage moved vertices since
  const i64 fcut =
      pass > 0 && !(O.balance_eps > 0.0f) ? cut : compute_cut(*L[0], m, part.p);
  HIP_TRY(hipMemcpyAsync(part_out, part.p, n * sizeof(i32), hipMemcpyDeviceToHost, st()));
  HIP_TRY(hipMemcpyAsync(pc_out, pc.p, k * sizeof(i32), hipMemcpyDeviceToHost, st()));
  C.sync();
  tr.push(TGP_TRACE_FINAL, 0, 0, 0, n, m, fcut, pass, O.check_level ? fnv_host(part_out, n) : 0);
  *cut_out = fcut;
  if (S) {
    S->levels = pass;
    S->passes = passes_total;
    S->coarsest_n = L[pass]->n;
    S->coarsest_npairs = npairs;
    S->match_rounds = match_rounds;
    S->sorts_skipped = sorts_skipped;
    const Level& l0 = *L[0];
    S->level0_form = l0.examined ? 8 | (l0.sorted ? 1 : 0) | (l0.symmetric ? 2 : 0) |
                                       (l0.symmetric && l0.unitw ? 4 : 0)
                                 : 0;
    S->seconds_initpart = initpart_s;
    S->seconds_total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    S->trace_len = tr.len;
  }
  return TGP_OK;
}

}  // namespace tgp

// ====================================================================== C ABI ==
using namespace tgp;

namespace {

int validate_host_graph(const tgp_graph* g) {
  if (g->n <= 0 || g->m <= 0 || !g->xadj || !g->adjncy || !g->adjwgt) return TGP_EINVAL;
  if (g->xadj[0] != 0 || g->xadj[g->n] != g->m) return TGP_EINVAL;
  for (i64 i = 0; i < g->n; i++)
    if (g->xadj[i + 1] < g->xadj[i]) return TGP_EINVAL;
  for (i64 i = 0; i < g->m; i++)
    if (g->adjncy[i] < 0 || g->adjncy[i] >= g->n) return TGP_EINVAL;
  return TGP_OK;
}

// Device scratch of the entry point that allocates nothing
// (tgp_partition_noalloc): every block its partitions use comes from here.  One
// copy per GPU, loaded with the code object and held until the process ends.
// Host code never passes `scratch` itself (that is a host-side stand-in, not a
// device address): the address each device holds it at comes from
// hipGetSymbolAddress at set-up.
static __device__ __align__(256) unsigned char scratch[TGP_NOALLOC_SCRATCH_BYTES];

// The no-allocation contexts, one per device: built at the first call on the
// device (with it current) around that device's copy of `scratch`.
struct NoallocDevice {
  Context* ctx = nullptr;
  void* scratch_ptr = nullptr;
  std::unique_ptr<i64[]> stage;  // host staging area (ordinary memory)
};
std::mutex g_noalloc_mu;
std::map<int, NoallocDevice> g_noalloc;

// One context per process, created on first use and intentionally never
// destroyed (its teardown would race the HIP runtime's own at exit).  Calls
// are serialised by a mutex: one partition at a time per process.
std::mutex g_ctx_mu;
Context* default_context() {
  static Context* ctx = new Context();
  return ctx;
}

// ------------------------------------------------------------ batch tuning --
// Settings of the batched engine: lanes per light vertex and the sort tile of
// the refinement launch, and each graph's heavy-row threshold by its mean
// degree (arcs / vertices).  Read from the table TGP_BATCH_TUNE names, else
// batch_tune.txt beside the library, else built in (4 lanes, 4096-entry tile,
// threshold 96).  TGP_BATCH_SETTING="lanes,heavy,tile" overrides all of it
// with one setting for every graph.  Every setting gives the same results.
struct BatchTuning {
  int g = 4, stile = 4096, heavy = 96;
  struct Cls {
    double lo, hi;
    int heavy;
  };
  std::vector<Cls> cls;
  std::string source = "built-in";
  int heavy_for(i64 n, i64 m) const {
    const double d = n > 0 ? (double)m / (double)n : 0.0;
    for (const Cls& c : cls)
      if (d >= c.lo && d < c.hi) return c.heavy;
    return heavy;
  }
};

bool valid_lanes(int g) { return g == 2 || g == 4 || g == 8; }
bool valid_tile(int t) { return t == 2048 || t == 4096 || t == 8192; }

std::string library_dir() {
  Dl_info info{};
  if (dladdr((void*)&tgp_batch_tuning, &info) && info.dli_fname) {
    const std::string p = info.dli_fname;
    const size_t sl = p.rfind('/');
    return sl == std::string::npos ? std::string(".") : p.substr(0, sl);
  }
  return ".";
}

// Table lines ('#' starts a comment):
//   lanes <2|4|8>
//   sort_tile <2048|4096|8192>
//   heavy_default <threshold>
//   heavy <mean degree from> <mean degree below> <threshold>
bool parse_tuning(const std::string& path, BatchTuning* t) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) {
    set_error("cannot open the batch tuning table " + path);
    return false;
  }
  char line[512];
  int ln = 0;
  std::string why;
  while (why.empty() && std::fgets(line, sizeof(line), f)) {
    ln++;
    if (char* h = std::strchr(line, '#')) *h = 0;
    char key[32] = {0};
    if (std::sscanf(line, "%31s", key) != 1) continue;  // blank
    const std::string k = key;
    int a = 0;
    double lo = 0, hi = 0;
    char extra[8];
    if (k == "lanes") {
      if (std::sscanf(line, "%*s %d %7s", &a, extra) != 1 || !valid_lanes(a)) why = "lanes takes 2, 4 or 8";
      else t->g = a;
    } else if (k == "sort_tile") {
      if (std::sscanf(line, "%*s %d %7s", &a, extra) != 1 || !valid_tile(a))
        why = "sort_tile takes 2048, 4096 or 8192";
      else t->stile = a;
    } else if (k == "heavy_default") {
      if (std::sscanf(line, "%*s %d %7s", &a, extra) != 1 || a < 0) why = "heavy_default takes a threshold >= 0";
      else t->heavy = a;
    } else if (k == "heavy") {
      if (std::sscanf(line, "%*s %lf %lf %d %7s", &lo, &hi, &a, extra) != 3 || !(lo < hi) ||
          lo < 0 || a < 0)
        why = "heavy takes <degree from> <degree below> <threshold >= 0> with from < below";
      else t->cls.push_back({lo, hi, a});
    } else {
      why = "unknown entry '" + k + "'";
    }
  }
  std::fclose(f);
  if (!why.empty()) {
    set_error("batch tuning table " + path + ", line " + std::to_string(ln) + ": " + why);
    return false;
  }
  return true;
}

int load_batch_tuning(BatchTuning* t) {
  *t = BatchTuning{};
  std::string path;
  const char* env = std::getenv("TGP_BATCH_TUNE");
  if (env && *env) {
    path = env;
  } else {
    const std::string p = library_dir() + "/batch_tune.txt";
    if (FILE* f = std::fopen(p.c_str(), "r")) {
      std::fclose(f);
      path = p;
    }
  }
  if (!path.empty()) {
    if (!parse_tuning(path, t)) return TGP_EINVAL;
    t->source = path;
  }
  const char* set = std::getenv("TGP_BATCH_SETTING");
  if (set && *set) {
    int g = 0, h = 0, st = 0;
    char extra[8];
    if (std::sscanf(set, "%d,%d,%d%7s", &g, &h, &st, extra) != 3 || !valid_lanes(g) ||
        !valid_tile(st) || h < 0) {
      set_error("TGP_BATCH_SETTING takes lanes,heavy,tile with lanes 2, 4 or 8, heavy >= 0 and "
                "tile 2048, 4096 or 8192");
      return TGP_EINVAL;
    }
    t->g = g;
    t->heavy = h;
    t->stile = st;
    t->cls.clear();
    t->source = "TGP_BATCH_SETTING";
  }
  return TGP_OK;
}

// Whether the batched engine takes the item (repaired coarsening, block
// initial partition, k <= BATCH_MAXK, no profiling or trace).
bool engine_takes(const tgp_batch_item& it) {
  return it.o.repair && it.o.initpart == tgp_initpart_block && !it.o.profile &&
         !it.o.pass_dump && !it.o.learned && !it.o.check_level &&
         (!it.stats || !it.stats->trace) && it.o.k >= 1 && it.o.k <= BATCH_MAXK &&
         it.g.n >= 2 && it.g.n <= INT_MAX && it.g.m > 0 && it.g.m <= INT_MAX && it.part &&
         it.part_count && (it.o.device_input || validate_host_graph(&it.g) == TGP_OK);
}

// The batched engine over a set of items: their graphs uploaded (when held
// in host memory), per-graph output slots, workgroup scratch and the level
// pool; then the coarsening and refinement launches, and the results.
struct BatchEngine {
  Context& C;
  tgp_batch_item* items;
  std::vector<i64> take;  // item index of each engine graph
  i64 T = 0;
  DBuf<u8> up;
  DBuf<i32> parts, pcs, status, order;
  DBuf<i64> cuts;
  DBuf<BatchGraph> dg;
  std::vector<BatchGraph> hg;
  std::vector<i64> part_off;
  std::vector<i32> ord;
  BatchBuffers B{};
  void* mem = nullptr;
  i64 budget = 0;

  BatchEngine(Context& Cx, tgp_batch_item* it) : C(Cx), items(it) {}
  ~BatchEngine() {
    if (mem) (void)hipFree(mem);
  }

  // false when there is no memory for the level pool
  bool setup(const BatchTuning& tun) {
    T = (i64)take.size();
    i64 maxn = 0, sum_n = 0, up_bytes = 0;
    for (i64 i : take) {
      const tgp_graph& g = items[i].g;
      maxn = std::max<i64>(maxn, g.n);
      sum_n += g.n;
      if (!items[i].o.device_input) up_bytes += 8 * (g.n + 1) + 8 * g.m + 64;
    }
    // inputs held in host memory go up once, into one buffer
    hg.assign((size_t)T, BatchGraph{});
    if (up_bytes) up = DBuf<u8>(&C.pool, (size_t)up_bytes);
    i64 uo = 0;
    parts = DBuf<i32>(&C.pool, (size_t)sum_n);
    pcs = DBuf<i32>(&C.pool, (size_t)T * BATCH_MAXK);
    status = DBuf<i32>(&C.pool, (size_t)T);
    order = DBuf<i32>(&C.pool, (size_t)T);
    cuts = DBuf<i64>(&C.pool, (size_t)T);
    dg = DBuf<BatchGraph>(&C.pool, (size_t)T);
    part_off.assign((size_t)T, 0);
    i64 po = 0;
    for (i64 j = 0; j < T; j++) {
      const tgp_batch_item& it = items[take[j]];
      BatchGraph& b = hg[j];
      b.n = it.g.n;
      b.m = it.g.m;
      if (it.o.device_input) {
        b.xadj = it.g.xadj;
        b.adj = it.g.adjncy;
        b.w = it.g.adjwgt;
      } else {
        u8* base = up.p + uo;
        HIP_TRY(hipMemcpyAsync(base, it.g.xadj, 8 * (b.n + 1), hipMemcpyHostToDevice, C.s));
        HIP_TRY(hipMemcpyAsync(base + 8 * (b.n + 1), it.g.adjncy, 4 * b.m,
                               hipMemcpyHostToDevice, C.s));
        HIP_TRY(hipMemcpyAsync(base + 8 * (b.n + 1) + 4 * b.m, it.g.adjwgt, 4 * b.m,
                               hipMemcpyHostToDevice, C.s));
        b.xadj = (const i64*)base;
        b.adj = (const i32*)(base + 8 * (b.n + 1));
        b.w = (const i32*)(base + 8 * (b.n + 1) + 4 * b.m);
        uo += 8 * (b.n + 1) + 8 * b.m + 64;
        uo = (uo + 255) & ~(i64)255;
      }
      b.k = it.o.k;
      b.wide = it.o.wide;
      b.target = it.o.coarsen_floor > 0 ? it.o.coarsen_floor : std::max<i64>(500, b.n / 500);
      part_off[j] = po;
      b.part_out = parts.p + po;
      po += b.n;
      b.pc_out = pcs.p + j * BATCH_MAXK;
      b.cut_out = cuts.p + j;
      b.status = status.p + j;
      b.heavy = tun.heavy_for(b.n, b.m);
      b.ticks = nullptr;
      // the balance bound of Run::balance_finest
      b.cap = it.o.balance_eps > 0.0f
                  ? (i64)std::floor((1.0 + (double)it.o.balance_eps) *
                                    (double)((b.n + b.k - 1) / b.k))
                  : 0;
    }
    // largest graphs first, so the long ones start early
    ord.resize((size_t)T);
    for (i64 j = 0; j < T; j++) ord[j] = (i32)j;
    std::stable_sort(ord.begin(), ord.end(), [&](i32 a, i32 b) { return hg[a].m > hg[b].m; });
    HIP_TRY(hipMemcpyAsync(dg.p, hg.data(), sizeof(BatchGraph) * T, hipMemcpyHostToDevice, C.s));
    // device-resident graphs are checked on the device first; a malformed one
    // stays out of the engine and the per-item path rejects it (TGP_EINVAL)
    std::vector<i32> devl;
    for (i64 j = 0; j < T; j++)
      if (items[take[j]].o.device_input) devl.push_back((i32)j);
    if (!devl.empty()) {
      DBuf<i32> dl(&C.pool, devl.size()), bad(&C.pool, (size_t)T);
      HIP_TRY(hipMemcpyAsync(dl.p, devl.data(), sizeof(i32) * devl.size(),
                             hipMemcpyHostToDevice, C.s));
      HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(i32) * T, C.s));
      launch_batch_csr_check(dg.p, dl.p, (i32)devl.size(), bad.p, C.s);
      std::vector<i32> hb((size_t)T);
      HIP_TRY(hipMemcpyAsync(hb.data(), bad.p, sizeof(i32) * T, hipMemcpyDeviceToHost, C.s));
      C.sync();
      ord.erase(std::remove_if(ord.begin(), ord.end(), [&](i32 j) { return hb[j] != 0; }),
                ord.end());
    }
    HIP_TRY(hipMemcpyAsync(order.p, ord.data(), sizeof(i32) * ord.size(), hipMemcpyHostToDevice,
                           C.s));
    launch_fill_i32(status.p, T, 1, C.s);
    // Coarsening: one workgroup per CU; refinement: two per CU; each with
    // scratch for the largest graph.  The coarse levels of the graphs share
    // a pool (up to 70% of the free memory); graphs that find it full wait
    // for the next round.
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device));
    cus = std::max(cus, 1);
    B.maxn = maxn;
    B.balance = 0;
    for (const BatchGraph& b : hg) B.balance |= b.cap > 0;
    B.stile = tun.stile;
    B.g = tun.g;
    B.blocks_c = (int)std::min<i64>(T, cus);
    B.blocks_r = (int)std::min<i64>(T, 2 * (i64)cus);
    const i64 bc = batch_coarsen_bytes(maxn), br = batch_refine_bytes(maxn);
    i64 want = 1 << 24;
    for (i64 i : take) want += 6 * (8 * items[i].g.m + 16 * (items[i].g.n + 1));
    size_t freeb = 0, totb = 0;
    HIP_TRY(hipMemGetInfo(&freeb, &totb));
    const i64 fixed = bc * B.blocks_c + br * B.blocks_r + batch_record_bytes() * T + 4096;
    B.pool_cap = std::min<i64>(want, (i64)((double)freeb * 0.7) - fixed);
    // test knob: a smaller level pool, so graphs find it full and go again
    // in later rounds
    const i64 pool_limit = env_i64("TGP_BATCH_POOL_BYTES", 0);
    if (pool_limit > 0) B.pool_cap = std::min<i64>(B.pool_cap, pool_limit);
    if (B.pool_cap <= 0 || hipMalloc(&mem, (size_t)(fixed + B.pool_cap)) != hipSuccess) {
      (void)hipGetLastError();
      mem = nullptr;
      return false;
    }
    u8* m8 = (u8*)mem;
    B.next = (u32*)m8;
    B.pool_used = (u64*)(m8 + 64);
    B.recs = m8 + 4096;
    B.scratch_c = B.recs + ((batch_record_bytes() * T + 255) & ~(i64)255);
    B.scratch_r = B.scratch_c + bc * B.blocks_c;
    B.pool = m8 + fixed;
    budget = C.clock_khz * 2000;  // 2 s per level's matching
    return true;
  }

  // the graphs' statuses after the launches
  std::vector<i32> statuses() {
    std::vector<i32> h((size_t)T);
    HIP_TRY(hipMemcpyAsync(h.data(), status.p, sizeof(i32) * T, hipMemcpyDeviceToHost, C.s));
    C.sync();
    return h;
  }

  // Coarsening and refinement rounds until every graph is done or handed
  // back (status 1); returns the number of rounds.
  int run(u64* bst) {
    std::vector<i32> hstat;
    std::vector<i32> queue = ord;
    int rounds = 0;
    while (!queue.empty()) {
      rounds++;
      HIP_TRY(hipMemsetAsync(mem, 0, 4096, C.s));  // queue counters, pool offset
      if (rounds > 1)
        HIP_TRY(hipMemcpyAsync(order.p, queue.data(), sizeof(i32) * queue.size(),
                               hipMemcpyHostToDevice, C.s));
      launch_batch_engine(dg.p, order.p, (i32)queue.size(), B, budget, bst, C.s);
      hstat = statuses();
      // graphs that found the pool full go again, with the pool to themselves
      std::vector<i32> again;
      for (i32 j : queue)
        if (hstat[j] == 3) again.push_back(j);
      for (i32 j : again) hstat[j] = 1;
      if (again.size() == queue.size()) break;  // no progress: the pool is too small
      queue.swap(again);
      if (!queue.empty()) {
        HIP_TRY(hipMemcpyAsync(status.p, hstat.data(), sizeof(i32) * T, hipMemcpyHostToDevice,
                               C.s));
        C.sync();
      }
    }
    HIP_TRY(hipMemcpyAsync(status.p, hstat.data(), sizeof(i32) * T, hipMemcpyHostToDevice, C.s));
    C.sync();
    return rounds;
  }

  // Results of the finished graphs into their items; the others are
  // appended to rest.
  void collect(std::vector<i64>& rest) {
    const std::vector<i32> hstat = statuses();
    std::vector<i32> hpc((size_t)T * BATCH_MAXK);
    std::vector<i64> hcut((size_t)T);
    HIP_TRY(hipMemcpyAsync(hpc.data(), pcs.p, sizeof(i32) * T * BATCH_MAXK,
                           hipMemcpyDeviceToHost, C.s));
    HIP_TRY(hipMemcpyAsync(hcut.data(), cuts.p, sizeof(i64) * T, hipMemcpyDeviceToHost, C.s));
    for (i64 j = 0; j < T; j++) {
      if (hstat[j] != 0) continue;
      tgp_batch_item& it = items[take[j]];
      HIP_TRY(hipMemcpyAsync(it.part, parts.p + part_off[j], sizeof(i32) * it.g.n,
                             hipMemcpyDeviceToHost, C.s));
    }
    C.sync();
    for (i64 j = 0; j < T; j++) {
      tgp_batch_item& it = items[take[j]];
      if (hstat[j] != 0) {
        rest.push_back(take[j]);
        continue;
      }
      std::memcpy(it.part_count, hpc.data() + j * BATCH_MAXK, sizeof(i32) * it.o.k);
      it.edge_cut = hcut[j];
      it.rc = TGP_OK;
      if (it.stats) {
        tgp_trace* tb = it.stats->trace;
        const i64 tc = it.stats->trace_cap;
        std::memset(it.stats, 0, sizeof(*it.stats));
        it.stats->trace = tb;
        it.stats->trace_cap = tc;
      }
    }
  }
};

// The batched engine over the items it takes; the others, and every graph
// the engine hands back, are appended to `rest`.
int engine_batch(Context& C, tgp_batch_item* items, i64 count, std::vector<i64>& rest) {
  BatchTuning tun;
  const int trc = load_batch_tuning(&tun);
  if (trc != TGP_OK) return trc;
  BatchEngine E(C, items);
  for (i64 i = 0; i < count; i++) {
    if (engine_takes(items[i])) E.take.push_back(i);
    else rest.push_back(i);
  }
  if (E.take.empty()) return TGP_OK;
  if (!E.setup(tun)) {
    for (i64 i : E.take) rest.push_back(i);
    return TGP_OK;
  }
  const bool trace = env_i64("TGP_BATCH_TRACE", 0) != 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DBuf<u64> bst;
  if (trace) {
    bst = DBuf<u64>(&C.pool, 13);
    HIP_TRY(hipMemsetAsync(bst.p, 0, 13 * sizeof(u64), C.s));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, C.s));
  }
  const int rounds = E.run(bst.p);
  if (trace) HIP_TRY(hipEventRecord(e1, C.s));
  const std::vector<i32> hstat = E.statuses();
  E.collect(rest);
  if (trace) {
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    i64 back = 0;
    for (i64 j = 0; j < E.T; j++) back += hstat[j] != 0;
    std::fprintf(stderr,
                 "# batch engine: %lld graphs, %lld handed back, %d + %d workgroups, "
                 "level pool %.1f GB, %d rounds, %.2f ms (lanes %d, sort tile %d, tuning %s)\n",
                 (long long)E.T, (long long)back, E.B.blocks_c, E.B.blocks_r,
                 E.B.pool_cap / 1e9, rounds, ms, E.B.g, E.B.stile, tun.source.c_str());
    u64 hb[13];
    HIP_TRY(hipMemcpy(hb, bst.p, sizeof(hb), hipMemcpyDeviceToHost));
    const char* nm[10] = {"check", "match", "contract", "gain", "token",
                          "moves", "sort",  "select",   "cut",  "misc"};
    std::fprintf(stderr, "# batch engine phases (workgroup-ms):");
    for (int q = 0; q < 10; q++)
      std::fprintf(stderr, " %s=%.1f", nm[q], hb[q] / (double)C.clock_khz);
    std::fprintf(stderr, " balance=%.1f passes=%llu levels=%llu\n",
                 hb[12] / (double)C.clock_khz, (unsigned long long)hb[10],
                 (unsigned long long)hb[11]);
  }
  return TGP_OK;
}

// The tuner: one coarsening launch over the sample, then the refinement
// launch again under each setting, timing every graph (wall-clock ticks of
// its workgroup) and checking that every setting gives the first one's
// results.  Settings: lanes {2, 4, 8} x heavy-row thresholds {48, 96, 192,
// 384, 768, 1536} at the 4096-entry tile, the threshold chosen per mean-degree class
// ([0, 12), [12, 20), [20, inf)); then the tiles {2048, 4096, 8192} at the
// chosen lanes and thresholds.  The sum of the graphs' workgroup times is
// the measure (the persistent launch's length follows it when the queue
// keeps every workgroup busy).
int autotune_on(Context& C, tgp_batch_item* items, i64 count, const char* path,
                std::string& rep) {
  BatchEngine E(C, items);
  for (i64 i = 0; i < count; i++)
    if (engine_takes(items[i])) E.take.push_back(i);
    else items[i].rc = TGP_EINVAL;
  if (E.take.empty()) {
    set_error("autotune: no graph within the batched engine's scope (repaired coarsening, "
              "block initial partition, k <= 16)");
    return TGP_EINVAL;
  }
  const BatchTuning base;
  if (!E.setup(base)) {
    set_error("autotune: no device memory for the level pool");
    return TGP_ENOMEM;
  }
  HIP_TRY(hipMemsetAsync(E.mem, 0, 4096, C.s));
  launch_batch_coarsen(E.dg.p, E.order.p, (i32)E.T, E.B, E.budget, nullptr, C.s);
  const std::vector<i32> st0 = E.statuses();
  static const double CLS_LO[3] = {0.0, 12.0, 20.0}, CLS_HI[3] = {12.0, 20.0, INFINITY};
  std::vector<int> cls((size_t)E.T, -1);
  i64 ncls[3] = {0, 0, 0}, ntuned = 0;
  for (i64 j = 0; j < E.T; j++) {
    if (st0[j] != 2) {
      items[E.take[j]].rc = TGP_EINVAL;  // outside the engine's scope, or the pool was full
      continue;
    }
    const double d = (double)E.hg[j].m / (double)E.hg[j].n;
    cls[j] = d < CLS_HI[0] ? 0 : d < CLS_HI[1] ? 1 : 2;
    ncls[cls[j]]++;
    ntuned++;
  }
  if (!ntuned) {
    set_error("autotune: no graph of the sample was coarsened by the engine");
    return TGP_EINVAL;
  }
  DBuf<i32> status0(&C.pool, (size_t)E.T);
  DBuf<u64> ticks(&C.pool, (size_t)E.T);
  HIP_TRY(hipMemcpyAsync(status0.p, E.status.p, sizeof(i32) * E.T, hipMemcpyDeviceToDevice, C.s));
  for (i64 j = 0; j < E.T; j++) E.hg[j].ticks = ticks.p + j;
  const i64 sum_n = E.part_off.back() + E.hg.back().n;
  std::vector<i32> ref_part, cur_part((size_t)sum_n), ref_pc, cur_pc((size_t)E.T * BATCH_MAXK);
  std::vector<i64> ref_cut, cur_cut((size_t)E.T);
  std::vector<u64> ht((size_t)E.T);
  int trials = 0, differ = 0;
  char line[512];
  // one refinement launch; per-class workgroup-ms, then the launch's ms
  auto trial = [&](int g, int tile, const int* heavy, double* out) {
    for (i64 j = 0; j < E.T; j++)
      if (cls[j] >= 0) E.hg[j].heavy = heavy[cls[j]];
    HIP_TRY(hipMemcpyAsync(E.dg.p, E.hg.data(), sizeof(BatchGraph) * E.T, hipMemcpyHostToDevice,
                           C.s));
    HIP_TRY(hipMemcpyAsync(E.status.p, status0.p, sizeof(i32) * E.T, hipMemcpyDeviceToDevice,
                           C.s));
    HIP_TRY(hipMemsetAsync(E.B.next + 1, 0, sizeof(u32), C.s));
    E.B.g = g;
    E.B.stile = tile;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, C.s));
    launch_batch_refine(E.dg.p, E.order.p, (i32)E.T, E.B, nullptr, C.s);
    HIP_TRY(hipEventRecord(e1, C.s));
    HIP_TRY(hipMemcpyAsync(ht.data(), ticks.p, sizeof(u64) * E.T, hipMemcpyDeviceToHost, C.s));
    HIP_TRY(hipMemcpyAsync(cur_part.data(), E.parts.p, sizeof(i32) * sum_n,
                           hipMemcpyDeviceToHost, C.s));
    HIP_TRY(hipMemcpyAsync(cur_pc.data(), E.pcs.p, sizeof(i32) * E.T * BATCH_MAXK,
                           hipMemcpyDeviceToHost, C.s));
    HIP_TRY(hipMemcpyAsync(cur_cut.data(), E.cuts.p, sizeof(i64) * E.T, hipMemcpyDeviceToHost,
                           C.s));
    C.sync();
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const std::vector<i32> st = E.statuses();
    for (int c = 0; c < 3; c++) out[c] = 0.0;
    bool same = true;
    for (i64 j = 0; j < E.T; j++) {
      if (cls[j] < 0) continue;
      if (st[j] != 0) same = false;
      out[cls[j]] += (double)ht[j] / (double)C.clock_khz;
    }
    out[3] = ms;
    if (ref_part.empty()) {
      ref_part = cur_part;
      ref_pc = cur_pc;
      ref_cut = cur_cut;
    } else {
      for (i64 j = 0; j < E.T && same; j++) {
        if (cls[j] < 0) continue;
        same = ref_cut[j] == cur_cut[j] &&
               !std::memcmp(&ref_part[E.part_off[j]], &cur_part[E.part_off[j]],
                            sizeof(i32) * E.hg[j].n) &&
               !std::memcmp(&ref_pc[j * BATCH_MAXK], &cur_pc[j * BATCH_MAXK],
                            sizeof(i32) * E.hg[j].k);
      }
    }
    trials++;
    differ += !same;
    std::snprintf(line, sizeof(line),
                  "lanes %d tile %4d heavy %4d/%4d/%4d: workgroup-ms %10.1f %10.1f %10.1f  "
                  "sum %10.1f  launch %8.1f ms%s\n",
                  g, tile, heavy[0], heavy[1], heavy[2], out[0], out[1], out[2],
                  out[0] + out[1] + out[2], out[3], same ? "" : "  RESULTS DIFFER");
    rep += line;
  };
  std::snprintf(line, sizeof(line),
                "# sample: %lld graphs tuned (mean degree [0,12): %lld, [12,20): %lld, >= 20: "
                "%lld)\n# per class: the sum of the graphs' workgroup-ms\n",
                (long long)ntuned, (long long)ncls[0], (long long)ncls[1], (long long)ncls[2]);
  rep += line;
  static const int LANES[3] = {2, 4, 8}, HEAVY[6] = {48, 96, 192, 384, 768, 1536};
  double best_total = INFINITY;
  int bg = 4, bh[3] = {96, 96, 96}, bu = 96;
  for (int gi = 0; gi < 3; gi++) {
    double cbest[3] = {INFINITY, INFINITY, INFINITY}, ubest = INFINITY;
    int ch[3] = {96, 96, 96}, uh = 96;
    for (int hi = 0; hi < 6; hi++) {
      const int h3[3] = {HEAVY[hi], HEAVY[hi], HEAVY[hi]};
      double r[4];
      trial(LANES[gi], 4096, h3, r);
      for (int c = 0; c < 3; c++)
        if (ncls[c] && r[c] < cbest[c]) {
          cbest[c] = r[c];
          ch[c] = HEAVY[hi];
        }
      if (r[0] + r[1] + r[2] < ubest) {
        ubest = r[0] + r[1] + r[2];
        uh = HEAVY[hi];
      }
    }
    double tot = 0;
    for (int c = 0; c < 3; c++)
      if (ncls[c]) tot += cbest[c];
    if (tot < best_total) {
      best_total = tot;
      bg = LANES[gi];
      for (int c = 0; c < 3; c++) bh[c] = ch[c];
      bu = uh;
    }
  }
  static const int TILES[3] = {2048, 4096, 8192};
  int bt = 4096;
  double bt_ms = INFINITY;
  for (int ti = 0; ti < 3; ti++) {
    double r[4];
    trial(bg, TILES[ti], bh, r);
    if (r[0] + r[1] + r[2] < bt_ms) {
      bt_ms = r[0] + r[1] + r[2];
      bt = TILES[ti];
    }
  }
  std::string table = "# batched engine tuning table, written by tgp_batch_autotune\n";
  {
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, C.device) == hipSuccess) {
      std::snprintf(line, sizeof(line), "# device: %s%s%s, %d CUs\n", prop.name,
                    prop.name[0] ? " " : "", prop.gcnArchName, prop.multiProcessorCount);
      table += line;
    }
  }
  std::snprintf(line, sizeof(line),
                "# sample: %lld graphs; %d settings measured, results %s\n"
                "lanes %d\nsort_tile %d\nheavy_default %d\n",
                (long long)ntuned, trials, differ ? "DIFFER" : "identical in every setting",
                bg, bt, bu);
  table += line;
  for (int c = 0; c < 3; c++) {
    if (!ncls[c]) continue;
    if (c < 2) std::snprintf(line, sizeof(line), "heavy %g %g %d\n", CLS_LO[c], CLS_HI[c], bh[c]);
    else std::snprintf(line, sizeof(line), "heavy %g inf %d\n", CLS_LO[c], bh[c]);
    table += line;
  }
  rep += table;
  if (differ) {
    set_error("autotune: results differ between engine settings in " + std::to_string(differ) +
              " of " + std::to_string(trials) + " trials");
    return TGP_EDIFF;
  }
  if (path && *path) {
    FILE* f = std::fopen(path, "w");
    if (!f || std::fputs(table.c_str(), f) < 0) {
      if (f) std::fclose(f);
      set_error(std::string("autotune: cannot write ") + path);
      return TGP_EIO;
    }
    std::fclose(f);
  }
  // the last trial's results (the chosen setting's) into the items
  std::vector<i64> rest;
  E.collect(rest);
  for (i64 i : rest) items[i].rc = TGP_EINVAL;
  return TGP_OK;
}

// Contexts of the batched form's workers, created on first use and, like the
// default context, never destroyed.
Context* worker_context(int w) {
  static std::vector<Context*> pool;
  while ((int)pool.size() <= w) pool.push_back(new Context());
  return pool[w];
}

// One partition on context C (its stream, pool and profile); the caller
// holds the lock that serialises use of C.
int partition_on(Context& Cx, const tgp_graph* g, const tgp_options* o, int32_t* part,
                 int32_t* part_count, int64_t* edge_cut, tgp_stats* stats) {
  last_error().clear();
  if (!g || !o || !part || !part_count || !edge_cut) {
    set_error("null argument");
    return TGP_EINVAL;
  }
  if (o->k <= 0 || o->k > 4096) {
    set_error("k must be in [1, 4096]");
    return TGP_EINVAL;
  }
  if (!o->device_input) {
    const int rc = validate_host_graph(g);
    if (rc != TGP_OK) {
      set_error("malformed CSR: need xadj[0]=0, non-decreasing offsets, xadj[n]=m>0, "
                "columns in [0,n)");
      return rc;
    }
  } else if (g->n <= 0 || g->m <= 0) {
    set_error("empty graph");
    return TGP_EINVAL;
  }
  if (o->balance_eps < 0.0f || o->balance_eps != o->balance_eps ||
      (o->balance_eps > 0.0f && o->k > 1024)) {
    set_error("balance_eps must be >= 0 (and k <= 1024 when it is set)");
    return TGP_EINVAL;
  }
  if (!o->wide) {  // reference int32 limits (gpu_part.cu:531, 886-888)
    if (g->n > INT_MAX || g->m > INT_MAX || g->n * (i64)o->k > (i64)INT_MAX) {
      set_error("n, m or n*k exceeds the reference's int32 limits; use wide mode");
      return TGP_ELIMIT;
    }
  }
  if (stats) {
    tgp_trace* tb = stats->trace;
    const i64 tc = stats->trace_cap;
    std::memset(stats, 0, sizeof(*stats));
    stats->trace = tb;
    stats->trace_cap = tc;
  }
  try {
    Context* C = &Cx;
    C->prof.on = o->profile != 0;
    for (int i = 0; i < KID_COUNT; i++) {
      C->prof.ms[i] = 0;
      C->prof.launches[i] = 0;
      C->prof.bytes[i] = 0;
    }
    int rc;
    {
      Run r(*C, *o, stats);
      rc = r.go(g, part, part_count, edge_cut);
    }
    C->sync();
    if (stats) {
      for (int i = 0; i < KID_COUNT && i < TGP_NKERNELS; i++) {
        stats->kernel_ms[i] = C->prof.ms[i];
        stats->kernel_launches[i] = C->prof.launches[i];
        stats->kernel_bytes[i] = C->prof.bytes[i];
      }
    }
    return rc;
  } catch (const Fail& f) {
    (void)hipStreamSynchronize(Cx.s);
    return f.code;
  }
}

// The balance stage alone (tgp_balance) on context C.
int balance_on(Context& Cx, const tgp_graph* g, const tgp_options* o, int32_t* part,
               int32_t* part_count, int64_t* edge_cut, tgp_stats* stats) {
  last_error().clear();
  if (!g || !o || !part || !part_count || !edge_cut) {
    set_error("null argument");
    return TGP_EINVAL;
  }
  if (o->k <= 0 || o->k > 1024 || !(o->balance_eps > 0.0f)) {
    set_error("tgp_balance needs k in [1, 1024] and balance_eps > 0");
    return TGP_EINVAL;
  }
  if (!o->device_input) {
    const int rc = validate_host_graph(g);
    if (rc != TGP_OK) {
      set_error("malformed CSR: need xadj[0]=0, non-decreasing offsets, xadj[n]=m>0, "
                "columns in [0,n)");
      return rc;
    }
  } else if (g->n <= 0 || g->m <= 0) {
    set_error("empty graph");
    return TGP_EINVAL;
  }
  if (stats) {
    tgp_trace* tb = stats->trace;
    const i64 tc = stats->trace_cap;
    std::memset(stats, 0, sizeof(*stats));
    stats->trace = tb;
    stats->trace_cap = tc;
  }
  try {
    Cx.prof.on = false;
    int rc;
    {
      Run r(Cx, *o, stats);
      rc = r.balance_only(g, part, part_count, edge_cut);
    }
    Cx.sync();
    return rc;
  } catch (const Fail& f) {
    (void)hipStreamSynchronize(Cx.s);
    return f.code;
  }
}

}  // namespace

extern "C" {

const char* tgp_last_error(void) { return last_error().c_str(); }

int tgp_batch_tuning(double mean_degree, int32_t* lanes, int32_t* heavy, int32_t* sort_tile) {
  last_error().clear();
  BatchTuning t;
  const int rc = load_batch_tuning(&t);
  if (rc != TGP_OK) return rc;
  const i64 scale = 1 << 20;
  if (lanes) *lanes = t.g;
  if (sort_tile) *sort_tile = t.stile;
  if (heavy) *heavy = t.heavy_for(scale, (i64)std::llround(mean_degree * (double)scale));
  return TGP_OK;
}

int tgp_batch_autotune(tgp_batch_item* items, int64_t count, const char* table_path,
                       char** report) {
  last_error().clear();
  if (report) *report = nullptr;
  if (count <= 0 || !items) {
    set_error("autotune: need items and count > 0");
    return TGP_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  std::string rep;
  int rc;
  try {
    rc = autotune_on(*default_context(), items, count, table_path, rep);
  } catch (const Fail& f) {
    (void)hipStreamSynchronize(default_context()->s);
    rc = f.code;
  }
  if (report) {
    *report = (char*)std::malloc(rep.size() + 1);
    if (*report) std::memcpy(*report, rep.c_str(), rep.size() + 1);
  }
  return rc;
}

int tgp_balance(const tgp_graph* g, const tgp_options* o, int32_t* part, int32_t* part_count,
                int64_t* edge_cut, tgp_stats* stats) {
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  return balance_on(*default_context(), g, o, part, part_count, edge_cut, stats);
}

int tgp_partition(const tgp_graph* g, const tgp_options* o, int32_t* part, int32_t* part_count,
                  int64_t* edge_cut, tgp_stats* stats) {
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  return partition_on(*default_context(), g, o, part, part_count, edge_cut, stats);
}

int tgp_partition_batch(tgp_batch_item* items, int64_t count, int32_t workers) {
  last_error().clear();
  if (count < 0 || (count > 0 && !items) || workers < 1 || workers > 64) {
    set_error("batch: need items, count >= 0 and workers in [1, 64]");
    return TGP_EINVAL;
  }
  if (count == 0) return TGP_OK;
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  // graphs within the engine's scope run in one persistent launch
  // (TGP_BATCH_ENGINE=0 sends every graph to the per-stream workers)
  std::vector<i64> rest;
  if (std::getenv("TGP_BATCH_ENGINE") && std::atoi(std::getenv("TGP_BATCH_ENGINE")) == 0) {
    for (i64 i = 0; i < count; i++) rest.push_back(i);
  } else {
    try {
      const int erc = engine_batch(*default_context(), items, count, rest);
      if (erc != TGP_OK) {
        for (i64 i = 0; i < count; i++) items[i].rc = erc;
        return erc;
      }
    } catch (const Fail& f) {
      (void)hipStreamSynchronize(default_context()->s);
      return f.code;
    }
  }
  const i64 nrest = (i64)rest.size();
  const int W = (int)std::min<i64>(workers, std::max<i64>(nrest, 1));
  std::vector<Context*> ctx;
  try {
    for (int w = 0; w < W; w++) ctx.push_back(worker_context(w));
  } catch (const Fail& f) {
    return f.code;
  }
  std::atomic<i64> next{0};
  auto work = [&](Context* C) {
    for (;;) {
      const i64 r = next.fetch_add(1);
      if (r >= nrest) break;
      tgp_batch_item& it = items[rest[r]];
      it.rc = partition_on(*C, &it.g, &it.o, it.part, it.part_count, &it.edge_cut, it.stats);
    }
  };
  std::vector<std::thread> th;
  for (int w = 1; w < W; w++) th.emplace_back(work, ctx[w]);
  work(ctx[0]);
  for (auto& t : th) t.join();
  for (i64 i = 0; i < count; i++)
    if (items[i].rc != TGP_OK) return items[i].rc;
  return TGP_OK;
}

const char* tgp_kernel_name(int id) {
  static const char* names[KID_COUNT] = {"match",  "relabel", "contract", "pairs", "project",
                                         "gain",   "token",   "compact",  "sort",  "select",
                                         "commit", "cut",     "scan",     "misc",
                                         "match_chain"};
  return (id >= 0 && id < KID_COUNT) ? names[id] : "";
}

int tgp_sort_moves(const float* gain, int64_t M, int64_t len, float* key, int32_t* idx,
                   void* stream) {
  if (!gain || !key || !idx || M < 0 || len < M || len > ((int64_t)1 << 31)) return TGP_EINVAL;
  if (len == 0) return TGP_OK;
  tgp::launch_sort_moves(gain, M, len, key, idx, (hipStream_t)stream);
  return hipGetLastError() == hipSuccess ? TGP_OK : TGP_EHIP;
}

int tgp_graph_upload(const tgp_graph* host, tgp_graph* dev) {
  if (!host || !dev || host->n <= 0 || host->m <= 0) return TGP_EINVAL;
  const int rc = validate_host_graph(host);
  if (rc != TGP_OK) return rc;
  void *x = nullptr, *a = nullptr, *w = nullptr;
  if (hipMalloc(&x, (host->n + 1) * sizeof(int64_t)) != hipSuccess ||
      hipMalloc(&a, host->m * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&w, host->m * sizeof(int32_t)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(x);
    (void)hipFree(a);
    (void)hipFree(w);
    return TGP_ENOMEM;
  }
  if (hipMemcpy(x, host->xadj, (host->n + 1) * sizeof(int64_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(a, host->adjncy, host->m * sizeof(int32_t), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(w, host->adjwgt, host->m * sizeof(int32_t), hipMemcpyHostToDevice) !=
          hipSuccess) {
    (void)hipFree(x);
    (void)hipFree(a);
    (void)hipFree(w);
    return TGP_EHIP;
  }
  dev->n = host->n;
  dev->m = host->m;
  dev->xadj = (const int64_t*)x;
  dev->adjncy = (const int32_t*)a;
  dev->adjwgt = (const int32_t*)w;
  return TGP_OK;
}

void tgp_graph_device_free(tgp_graph* dev) {
  if (!dev) return;
  (void)hipFree((void*)dev->xadj);
  (void)hipFree((void*)dev->adjncy);
  (void)hipFree((void*)dev->adjwgt);
  dev->xadj = nullptr;
  dev->adjncy = nullptr;
  dev->adjwgt = nullptr;
}

int tgp_partition_noalloc(const tgp_graph* g, const tgp_options* o, int32_t* part,
                          int32_t* part_count, int64_t* edge_cut, tgp_stats* stats) {
  last_error().clear();
  if (!g || !o) {
    set_error("null argument");
    return TGP_EINVAL;
  }
  if (g->n > TGP_NOALLOC_MAX_N || g->m > TGP_NOALLOC_MAX_M || o->k > TGP_NOALLOC_MAX_K) {
    set_error("tgp_partition_noalloc takes n <= " + std::to_string(TGP_NOALLOC_MAX_N) +
              ", m <= " + std::to_string(TGP_NOALLOC_MAX_M) + " and k <= " +
              std::to_string(TGP_NOALLOC_MAX_K) + " (its device scratch holds " +
              std::to_string(TGP_NOALLOC_SCRATCH_BYTES >> 20) +
              " MiB); use tgp_partition for larger graphs");
    return TGP_EINVAL;
  }
  if (o->learned || o->profile) {
    set_error("tgp_partition_noalloc runs neither the learned proposer nor profiling");
    return TGP_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    set_error("no HIP device");
    return TGP_EHIP;
  }
  NoallocDevice& nd = g_noalloc[dev];
  if (!nd.ctx) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      set_error(std::string("hipGetSymbolAddress of the scratch array failed: ") +
                hipGetErrorString(e));
      return TGP_EHIP;
    }
    nd.scratch_ptr = ptr;
    nd.stage.reset(new i64[Context::PINNED_ELEMS]);
    try {
      nd.ctx = new Context(nd.scratch_ptr, (size_t)TGP_NOALLOC_SCRATCH_BYTES, nd.stage.get());
    } catch (const Fail& f) {
      return f.code;
    }
  }
  return partition_on(*nd.ctx, g, o, part, part_count, edge_cut, stats);
}

void tgp_release_device_memory(void) {
  std::lock_guard<std::mutex> lock(g_ctx_mu);
  Context* C = default_context();
  C->sync();
  C->pool.trim();
}

}  // extern "C"
