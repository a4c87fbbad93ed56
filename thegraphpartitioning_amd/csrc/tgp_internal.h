// tgp_internal.h -- internal interface between the host driver and the gfx950
// kernels of the partitioner.  Not part of the public C ABI (include/tgp.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace tgp {

// Detail of the last failure on the calling thread (tgp_last_error).
inline std::string& last_error() {
  static thread_local std::string s;
  return s;
}

typedef int64_t i64;
typedef int32_t i32;
typedef uint32_t u32;
typedef uint64_t u64;
typedef uint8_t u8;
typedef uint16_t u16;

// A CSR level resident in HBM: 64-bit row offsets, 32-bit ids and weights.
struct DGraph {
  i64 n;          // vertices
  i64 m;          // arcs = xadj[n]
  const i64* xadj;
  const i32* adj;
  const i32* w;
};

// Kernel families, for per-kernel timing (tgp_kernel_name).
enum KernelId {
  KID_MATCH = 0,   // exact heavy-edge matching rounds
  KID_RELABEL,     // root scan + coarse ids + trace-back
  KID_CONTRACT,    // coarse CSR construction
  KID_PAIRS,       // unique undirected pair count of the coarsest graph
  KID_PROJECT,     // projection + part-count histogram
  KID_GAIN,        // gains + targets + candidates
  KID_TOKEN,       // token passing + move flags
  KID_COMPACT,     // move buffer compaction (scan + scatter)
  KID_SORT,        // bitonic network
  KID_SELECT,      // per-destination rank selection
  KID_COMMIT,      // commit of selected moves
  KID_CUT,         // edge cut
  KID_SCAN,        // generic scans
  KID_MISC,        // fills, tables, capacities
  KID_CHAIN,       // the chained matching kernel alone (the rest of matching is KID_MATCH)
  KID_COUNT
};

// Scratch provider for multi-level scans (device memory owned by the caller).
struct ScanScratch {
  i64* buf;       // at least scan_scratch_elems(n) elements
  i64 cap;
};
i64 scan_scratch_elems(i64 n);

// ---- generic -----------------------------------------------------------------
void launch_fill_i32(i32* p, i64 n, i32 v, hipStream_t s);
void launch_fill_i64(i64* p, i64 n, i64 v, hipStream_t s);
void launch_fill_u8(u8* p, i64 n, u8 v, hipStream_t s);
void launch_fill_u64(u64* p, i64 n, u64 v, hipStream_t s);
// out[i] = sum(in[0..i)), *total = sum(in); out may alias nothing.
void launch_scan_excl(const i64* in, i64* out, i64 n, i64* total, ScanScratch sc,
                      hipStream_t s);
// Row-degree census of a level: unit-weight flag, heavy list (range length >
// heavy_deg, ranges end at `nedges` for the last vertex), and *maxabs =
// max |w| over the level's arcs (zeroed by the caller).
void launch_level_census(DGraph g, i64 nedges, i64 heavy_deg, i32* heavy_list,
                         i32* heavy_count, i32* nonunit_flag, u64* maxabs, hipStream_t s);

// ---- coarsening ---------------------------------------------------------------
// Exact matching in rounds over the id window [lo, lo + W) (lo = smallest
// undecided id, kept in *lo on the device; the matching is done when *lo == n).
void launch_match_init(i64 n, i32* mate, u8* dec, i32* pend, u64* R, i64* lo, hipStream_t s);
void launch_match_round(DGraph g, i32* mate, u8* dec, i32* pend, u64* R, i32* prop, i64* lo,
                        i64 W, u32 round, hipStream_t s);
// Same matching for levels whose arcs are symmetric (cached blockers, no atomics).
void launch_match_sym_init(i64 n, i32* mate, u8* dec, i32* blkA, i32* blkC, i32* prop, i64* lo,
                           hipStream_t s);
// req/nreq: queue of deferred long-row scans (W pairs) and its two counters,
// both zeroed by the caller before the first round.
// Chained matching (symmetric levels) on the state st (-1 undecided, -2
// decided unmatched, else the partner): one launch decides vertices from
// lo[0] upward until all are decided or `budget` wall-clock ticks pass, then
// sets lo[0] to the first undecided vertex (n when done).  *ticket must be 0.
// stats (optional, 12 zeroed u64): wave iterations, idle iterations, short
// A/proposal scans, long A scans, long proposal scans, proposal-row fetches,
// short C scans, long C scans, decisions, entries of the long A, proposal and
// C scans.
// low (optional, off == nullptr for none): the lower-neighbour lists of the
// long rows, which the long-row scans of the "no undecided neighbour below"
// conditions read instead of the whole row.
struct ChainLow {
  const i64* off;  // n + 1: entries of v at adj[off[v], off[v+1]) (empty for short rows)
  const i32* adj;  // the entries u < v of row v, in row order
  int sorted = 0;  // every row is strictly ascending (so are the lists above)
  int unitw = 0;   // every arc weight is 1 (the heaviest free neighbour is the first one)
  // Best-entry lists (optional): row v with lidx[v] >= 0 has its lcnt[lidx[v]]
  // best entries (weight, then first position; v itself, entries outside
  // [0, n) and negative weights left out) at lists + 64 * lidx[v], best first,
  // and every other entry of the row ranks below them.
  const i32* lidx = nullptr;
  const i32* lists = nullptr;
  const i32* lcnt = nullptr;
};
constexpr int CHAIN_LIST_LEN = 64;
// Best-entry lists of the rows longer than minlen: lidx (n ints) = -1 or the
// row's list index, lrows[index] = the row, *nl (zeroed by the caller) = the
// count; lists (CHAIN_LIST_LEN ints per list) and lcnt receive the lists.
// At most m / minlen + 1 rows qualify.
void launch_chain_lists(DGraph g, i64 minlen, i32* lidx, i32* lrows, i32* nl, i32* lists,
                        i32* lcnt, hipStream_t s);
void launch_match_chain(DGraph g, ChainLow low, i32* st, i32* blkA, i32* blkC, i32* prop,
                        const i64* lo, u32* ticket, i64 budget, int blocks, int sg, int per,
                        u64* stats, hipStream_t s);
// Lower-neighbour lists of the rows longer than minlen: list/nlist (n ints,
// *nlist zeroed by the caller) receive those rows, cnt[v] their entry counts
// (cnt must be zeroed by the caller), then, with off = exclusive scan of cnt,
// launch_chain_low_write puts the entries into out.
void launch_chain_low_count(DGraph g, i64 minlen, i32* list, i32* nlist, i64* cnt,
                            hipStream_t s);
void launch_chain_low_write(DGraph g, const i32* list, const i32* nlist, const i64* off, i32* out,
                            hipStream_t s);
// *out = the level's longest row (xadj[v+1] - xadj[v]); *out zeroed by the caller
void launch_max_row(DGraph g, u64* out, hipStream_t s);
// lo[0] = first vertex at or above lo[0] with st == -1 (n if none); lo[3]
// must hold n on entry and holds n again on exit
void launch_first_undecided(const i32* st, i64 n, i64* lo, hipStream_t s);
// mate = st where st >= 0 else -1; dec = (st != -1)
void launch_chain_export(const i32* st, i64 n, i32* mate, u8* dec, hipStream_t s);

// ---- batched engine: one workgroup per graph, the whole partition on device
// One graph of a batch (device pointers).  status on return: 0 done, 1 the
// graph needs the multi-kernel path (outside the engine's limits); 2 marks a
// coarsened graph between the two launches.
struct BatchGraph {
  i64 n, m;
  const i64* xadj;
  const i32* adj;
  const i32* w;
  i32 k, wide;
  i64 target;  // coarsening stops at or below this many vertices
  i32* part_out;
  i32* pc_out;
  i64* cut_out;
  i32* status;
  i32 heavy;   // rows longer than this are refined by a whole wave
  u64* ticks;  // optional: the refinement's wall-clock ticks for this graph
  i64 cap;     // > 0: the balance stage's bound (tgp_options.balance_eps), 0: none
};
constexpr int BATCH_MAXK = 16;
// sort tiles the refinement launch is built for (entries held in LDS)
constexpr int BATCH_STILES[3] = {2048, 4096, 8192};
// lanes per light vertex it is built for
constexpr int BATCH_LANES[3] = {2, 4, 8};
// Scratch of one coarsening / refinement workgroup for graphs of at most
// maxn vertices, and the size of one graph's level record.
i64 batch_coarsen_bytes(i64 maxn);
i64 batch_refine_bytes(i64 maxn);
i64 batch_record_bytes();
struct BatchBuffers {
  i64 maxn;
  u8* scratch_c;  // blocks_c x batch_coarsen_bytes(maxn)
  int blocks_c;
  u8* scratch_r;  // blocks_r x batch_refine_bytes(maxn)
  int blocks_r;
  u8* pool;       // coarse levels of every graph
  i64 pool_cap;
  u64* pool_used; // zero before the launch
  u8* recs;       // count x batch_record_bytes()
  u32* next;      // three queue counters, zero before the launches
  int stile;      // the refinement launch's sort tile (one of BATCH_STILES)
  int g;          // its lanes per light vertex (one of BATCH_LANES)
  int balance;    // some graph has a balance bound: the balance launch runs
};
// The coarsening launch over the queue, then the refinement launch, then
// (B.balance) the balance launch.
// stats (optional, 13 u64): wall-clock ticks per phase summed over
// workgroups (check, match, contract, gain, token, moves, sort, select, cut,
// misc), refinement passes, levels built, then balance-stage ticks.
void launch_batch_engine(const BatchGraph* graphs, const i32* order, i32 count,
                         const BatchBuffers& B, i64 budget, u64* stats, hipStream_t s);
// The balance stage's launch (graphs with cap > 0 and status 0; next[2]).
void launch_batch_balance(const BatchGraph* graphs, const i32* order, i32 count,
                          const BatchBuffers& B, u64* stats, hipStream_t s);
// The two launches apart (the tuner coarsens once and refines many times;
// the refinement takes the graphs whose status is 2 and reads next[1]).
void launch_batch_coarsen(const BatchGraph* graphs, const i32* order, i32 count,
                          const BatchBuffers& B, i64 budget, u64* stats, hipStream_t s);
void launch_batch_refine(const BatchGraph* graphs, const i32* order, i32 count,
                         const BatchBuffers& B, u64* stats, hipStream_t s);
void launch_match_sym_round(DGraph g, i32* mate, u8* dec, i32* blkA, i32* blkC, i32* prop,
                            i32* decision, i64* lo, i64 W, i32* req, i32* nreq, u32 round,
                            hipStream_t s);
// Arc checks of a level (table: cap >= 2m slots, pow2; wt: cap ints):
// flags[0] |= some arc s->x (s != x, x < n) has no reverse; flags[1] |= some
// (s, x) appears twice in s's row; flags[2] |= some reverse has another weight.
void launch_sym_check(DGraph g, u64* table, i32* wt, i64 cap, i32* flags, hipStream_t s);
// *flag = 1 (zeroed by the caller) unless xadj[0] = 0, the offsets do not
// decrease, xadj[n] = m and every adj entry is in [0, n) (a CSR handed over
// in device memory)
void launch_csr_check(const i64* xadj, const i32* adj, i64 n, i64 m, i32* flag, hipStream_t s);
// bad[j] |= 1 (bad zeroed by the caller) for each listed batch graph j whose
// device CSR fails the same check
void launch_batch_csr_check(const BatchGraph* dg, const i32* list, i32 cnt, i32* bad,
                            hipStream_t s);
// flags[3] |= some row is not strictly ascending; flags[5] |= some arc between
// distinct vertices in range weighs other than 1.  crow (sym_check_chunks(m)
// entries, m = xadj[n]) receives the row of each chunk's first arc, which
// launch_sym_check_sorted then reads.
void launch_rows_strict(DGraph g, i64 m, i32* crow, i32* flags, hipStream_t s);
// the same checks for strictly ascending rows, by binary search of the
// reverse arc in its (sorted) row: flags[0], flags[2] as above; of each pair
// the arc into the row of lower (degree bit length, id) searches: cnt[1] =
// searching arcs whose reverse was found, cnt[0] = the other arcs (symmetric
// iff flags[0] == 0 and they agree); flags[4] |= a found pair's weight is not
// 1; cls (n bytes) receives the degree bit lengths; crow as launch_rows_strict
// left it
i64 sym_check_chunks(i64 m);
void launch_sym_check_sorted(DGraph g, i64 m, u8* cls, const i32* crow, i32* flags, u64* cnt,
                             hipStream_t s);
void launch_relabel(const i32* mate, i64 n, i64* flags, i64* pos, i64* total,
                    i32* new_ids, i32* tb, ScanScratch sc, hipStream_t s);
// Reference contraction (gpu_part.cu:145-242 semantics): padded rows of 64.
void launch_contract_ref(DGraph g, const i32* mate, const i32* new_ids, i64 n_c,
                         i32* sadj, i32* sw, i64* cnt, hipStream_t s);
void launch_compact_rows(const i32* sadj, const i32* sw, const i64* ox, i64 n_c,
                         i32* oa, i32* ow, hipStream_t s);
constexpr i64 REP_CHUNK_ENTRIES = 4096;
// Repaired contraction.  Coarse row c enumerates fine rows tb[2c] then
// tb[2c+1] (d entries); classify writes d, dlong = d for rows past 64 entries
// (their entry words R sit at the exclusive scan eoff of dlong), and for rows
// past REP_MED entries (long rows, listed in lng) their global table capacity.
// (nchk[e] = chunks of long row e)
// (medium rows: med holds those of more than 256 entries, meds the others)
void launch_rep_classify(const i64* xadj, const i32* tb, i64 n_c, i64* dsz, i64* dlong,
                         i64* csz, i32* med, i32* nmed, i32* lng, i32* nlng, i64* nchk,
                         i32* meds, i32* nmeds, hipStream_t s);
// rows of at most REP_MED entries (registers, or an LDS table for the listed
// medium rows, whose entry words go to R at eoff): pass 0 writes cnt[c], pass
// 1 writes row c at ox[c]
void launch_rep_rows(DGraph g, const i32* tb, const i32* new_ids, i64 n_c, const i64* dsz,
                     const i32* med, i64 nmed, const i32* meds, i64 nmeds, const i64* eoff,
                     u64* R, int pass, i64* cnt, const i64* ox, i32* oa, i32* ow,
                     hipStream_t s);
// Long rows: listed in lng (nl rows), cut into chunks of REP_CHUNK_ENTRIES
// (cbase[e] = first chunk of row e, exclusive scan; ch_row/ch_beg describe
// each chunk); per-row global tables T (64-bit slots) and W at toff with
// capacity csz; first-occurrence slots E at eoff; per-chunk counts and their
// exclusive scan choff.
struct RepLong {
  const i32* lng = nullptr;
  i64 nl = 0;
  const i64* dsz = nullptr;
  const i64* cbase = nullptr;
  const i32* ch_row = nullptr;
  const i64* ch_beg = nullptr;
  i64 nch = 0;
  const i64* toff = nullptr;
  const i64* csz = nullptr;
  u64* T = nullptr;
  i32* W = nullptr;
  const i64* eoff = nullptr;
  i32* E = nullptr;
  i64* chunk_cnt = nullptr;
  const i64* choff = nullptr;
};
void launch_rep_chunks(const i32* lng, i64 nl, const i64* dsz, const i64* cbase, i32* ch_row,
                       i64* ch_beg, hipStream_t s);
// pass 0: insert + mark (chunk counts); pass 1: emit at ox
void launch_rep_long(DGraph g, const i32* tb, const i32* new_ids, const RepLong& L, int pass,
                     const i64* ox, i32* oa, i32* ow, hipStream_t s);
// cnt[c] of the long rows from the scanned chunk counts
void launch_rep_rowcnt(const RepLong& L, i64* cnt, hipStream_t s);
// Long rows, bucketed form (the default; the global tables are the fallback).
// Entry words R (x << 32 | w in position order, at eoff); per long row e,
// 2^lgb[e] buckets (nbk[e]) from bucket bbase[e] on (exclusive scan, nbt in
// all); brow = row of each bucket; bcnt = entries per bucket (zeroed before
// fetch), boff = their exclusive scan, bcur = fill cursors (zeroed); K / W =
// bucket records (x << 32 | position, weight); over = set when a bucket's
// distinct neighbours pass the fill limit (the level then takes the tables).
struct RepBuckets {
  i32* lgb = nullptr;
  i64* nbk = nullptr;
  const i64* bbase = nullptr;
  i64 nbt = 0;
  i32* brow = nullptr;
  u64* bcnt = nullptr;
  const i64* boff = nullptr;
  u32* bcur = nullptr;
  u64* R = nullptr;
  u64* K = nullptr;
  i32* W = nullptr;
  i32* over = nullptr;
};
// most distinct neighbours one bucket's LDS table takes before the level
// falls back to the global tables
constexpr i32 REP_BUCKET_FILL = 3072;
// target: entries per bucket aimed at (0: the built-in 1024)
void launch_rl_buckets(const RepLong& L, i64 target, i32* lgb, i64* nbk, hipStream_t s);
void launch_rl_brow(i64 nl, const i64* bbase, i32* brow, hipStream_t s);
void launch_rl_fetch(DGraph g, const i32* tb, const i32* new_ids, const RepLong& L,
                     const RepBuckets& B, hipStream_t s);
void launch_rl_scatter(const RepLong& L, const RepBuckets& B, hipStream_t s);
void launch_rl_dedup(const RepLong& L, const RepBuckets& B, i32 fill, hipStream_t s);
// chunk_cnt of every chunk from R
void launch_rl_count(const RepLong& L, const RepBuckets& B, hipStream_t s);
void launch_rl_emit(const RepLong& L, const RepBuckets& B, const i64* ox, i32* oa, i32* ow,
                    hipStream_t s);
// Unique undirected pairs (min,max), u != v, 0 <= u < n, true row ends.
void launch_pairs(DGraph g, u64* table, i64 cap, u64* count, hipStream_t s);
// count += arcs v -> u with 0 <= u < n, u != v (same row ranges as
// launch_pairs); half of it is the unique pair count on a symmetric level
// whose rows hold no repeats
void launch_count_pair_arcs(DGraph g, u64* count, hipStream_t s);

// ---- projection / counts ------------------------------------------------------
void launch_project(const i32* tb, const i32* part_c, i64 n_c, i32* part_f, i64 n_f,
                    i32 k, hipStream_t s);
void launch_sanitize_hist(i32* part, i64 n, i32 k, i32* pc, hipStream_t s);
void launch_block_init(i32* part, i64 n, i32 k, hipStream_t s);

// ---- refinement ---------------------------------------------------------------
// per-pass balance table bg[2][k] and src_big[k]; also zeroes
// zero_k[0..nzero) and zero_2[0..1] and sets *one1 = 1 when given
void launch_bg_table(const i32* pc, i32 k, i64 n, int phase, float* bg, u8* src_big,
                     i32* zero_k, i32 nzero, i32* zero_2, i32* one1, hipStream_t s);
// Heavy rows (> 64 arcs) split into medium (<= 4096) and huge lists.
void launch_split_heavy(DGraph g, i64 nedges, const i32* heavy, i64 nheavy, i32* mid,
                        i32* nmid, i32* medium, i32* nmedium, i32* huge, i32* nhuge,
                        hipStream_t s);
// Gains: 8 lanes per light row, a wave per medium row, a block per huge row
// (k <= 512); one wave per vertex for larger k.
// When cut_acc (accumulator slots) is given and k <= 512, also adds the edge
// cut of `part` (the value launch_cut_arcs would add) to it and returns true; otherwise
// returns false and the caller computes the cut separately.
// Row classes of a level for the gain kernels: mid rows (64 < len <= 512),
// medium rows (512 < len <= 4096),
// long rows (> 4096) and the long rows' chunks of at most GAIN_CHUNK_ARCS arcs
// with per-row accumulators (nhuge * k tallies, nhuge |w| sums, all zero
// between passes).
constexpr i64 GAIN_CHUNK_ARCS = 4096;
struct GainLists {
  const i32* mid = nullptr;
  i64 nmid = 0;
  const i32* medium = nullptr;
  i64 nmedium = 0;
  const i32* huge = nullptr;
  i64 nhuge = 0;
  const i32* ch_row = nullptr;
  const i64* ch_beg = nullptr;
  const i64* ch_end = nullptr;
  i64 nch = 0;
  i32* gtally = nullptr;
  i64* gabs = nullptr;
  int all_exact = 0;  // every row's sum of |w| is at most 2^24 (integer tallies are exact)
  u32* ctr = nullptr;  // 8 work counters (256 u32) of the fused gain kernel, reset by it
};
// over[0] += rows whose sum of |w| passes 2^24, listed at over + 1 (room for
// n); heavy = rows longer than heavy_deg; rows with degree * *maxabs <= 2^24
// (the census's max |w|) are not summed
void launch_row_abs(DGraph g, i64 nedges, i64 heavy_deg, const i32* heavy, i64 nheavy,
                    const u64* maxabs, i32* over, hipStream_t s);
// part8 (optional, k <= PART8_MAXK): byte copy of part that the neighbour
// gathers read instead of part.
constexpr i32 PART8_MAXK = 255;
// Kept tallies (T[v][q] = weight of v's arcs into part q).  A row whose arcs
// all weigh >= 0 and sum to at most 255 (and that is not the last row) keeps
// its tallies as bytes: t8 + v * tally8_stride(k) (k rounded up to 16), marked
// bix[v] < 0; every other row keeps ints at t32 + bix[v] * tally_row_stride(k)
// (k rounded up to 4).  Rows are 16-byte aligned.  A byte tally never leaves
// [0, 255], so byte updates are exact 32-bit atomic adds on the word holding
// the byte (no carry or borrow crosses a byte).
// With all_int set every row is an int row at t32 + v * tally_row_stride(k)
// (no bix); a level where most rows need ints takes that layout.  bigmask has
// bit v set for the int rows of a mixed level (one bit per vertex, so the
// class test of a row is an L2 hit while its byte row is already in flight).
// bigv[b] (b < nbig) is the vertex of int row b on a mixed level.  Entries of
// a row past k (its padding) hold zero.  With t16 set (levels with k > 64)
// the rows that are not int rows are 16-bit rows at t16 + v * tally16_stride(k)
// (k rounded up to 8) instead of byte rows: arcs weigh >= 0 and sum to at
// most 65535, and the updates are exact 32-bit adds on the word holding the
// half.
struct TallyRef {
  u8* t8 = nullptr;
  i32* t32 = nullptr;
  const i32* bix = nullptr;
  const u32* bigmask = nullptr;
  int all_int = 0;
  const i32* bigv = nullptr;
  i32 nbig = 0;
  u16* t16 = nullptr;
};
inline i64 tally_row_stride(i32 k) { return (i64)((k + 3) & ~3); }
inline i64 tally8_stride(i32 k) { return (i64)((k + 15) & ~15); }
inline i64 tally16_stride(i32 k) { return (i64)((k + 7) & ~7); }
// short rows serve levels with 64 < k <= TALLY16_MAXK
constexpr i32 TALLY16_MAXK = 256;
// bix[v] = -1 for a byte or short row (arcs >= 0 summing to at most limit:
// 255 or 65535), else its int row (wave-aggregated from *nbig, zeroed by the
// caller), bigmask[(n + 31) / 32] the int rows; all_int puts every row in the
// int table.
void launch_tally_class(DGraph g, i64 nedges, int unit, int all_int, i64 limit, i32* bix,
                        u32* bigmask, i32* nbig, hipStream_t s);
// bigv[bix[v]] = v for the int rows of a mixed level
void launch_big_list(const i32* bix, i64 n, i32* bigv, hipStream_t s);
// Tout (optional): also writes each vertex's integer tally.
bool launch_gain(DGraph g, i64 nedges, const i32* part, const u8* part8, i32 k, const float* bg,
                 const u8* src_big, float* gain, i32* target, float* cand, int phase, int unit,
                 i32* gscratch, const GainLists& L, u64* cut_acc, const TallyRef* Tout,
                 hipStream_t s);
// Gains of the listed rows only, gathered from their arcs (k <= 512): rows
// whose sum of |w| passes 2^24 are tallied in fp32 in row order.
void launch_gain_rows(DGraph g, i64 nedges, const i32* part, i32 k, const float* bg,
                      const u8* src_big, float* gain, i32* target, float* cand, int phase,
                      int unit, const i32* list, i64 cnt, hipStream_t s);
// Gains (and the cut into cut_acc slots) from kept tallies T; lanes (2, 4, 8 or
// 16) read each row.
void launch_gain_tallies(i64 n, const i32* part, const TallyRef& T, i32 k, const float* bg,
                         const u8* src_big, float* gain, i32* target, float* cand, int phase,
                         u64* cut_acc, int lanes, hipStream_t s);
// T update for the applied moves app[0..*napp) (at most max_app) of a commit;
// skip_last leaves T[n-1] alone.
void launch_tally_update(DGraph g, const TallyRef& T, i32 k, const i32* app, const i32* napp,
                         i64 max_app, const i32* mv_v, const i32* mv_src, const i32* mv_dst,
                         int unit, int skip_last, hipStream_t s);
// T[n-1] from its range [xadj[n-1], nedges)
// Edge cut of part from kept tallies (sum over v of T[v][q] for q != part[v]),
// added into the accumulator slots.
void launch_tally_cut(i64 n, const i32* part, const TallyRef& T, i32 k, u64* cut_acc,
                      hipStream_t s);
void launch_tally_last(DGraph g, i64 nedges, const i32* part, const TallyRef& T, i32 k, int unit,
                       hipStream_t s);
// part8[v] = part[v] if 0 <= part[v] < k, else 255
void launch_part8(const i32* part, i64 n, i32 k, u8* part8, hipStream_t s);
// hb[2h], hb[2h+1] = arc range of long row h (the last row ends at nedges)
void launch_huge_bounds(DGraph g, i64 nedges, const i32* huge, i64 nhuge, i64* hb,
                        hipStream_t s);
i64 gain_gscratch_elems(i32 k);
// cand[v] = gain of v when v is a move candidate, NaN otherwise (the gain
// kernels write it); rows of at most heavy_deg arcs: lanes = 0 streams the
// arcs of 64 consecutive rows per wave, 1, 4 or 8 lanes per row otherwise.
// wit (optional, n entries, -1 or a neighbour of v, kept across the passes of
// a level): a neighbour that beat v in an earlier pass, checked before v's row
// is read and rewritten when the row finds another (the arc stream, one
// thread per row and the heavy rows use it; 4 or 8 lanes per row do not)
void launch_token(DGraph g, i64 nedges, const float* cand, const i32* part, const i32* target,
                  i32 k, u8* flag, const i32* heavy_list, i64 heavy_count, i64 heavy_deg,
                  int lanes, i32* wit, hipStream_t s);
// Winners (flag[v] = 1) to the move arrays in ascending vertex order; *total
// = their count, outc[src] += moves out of each part and (when given)
// inc[dst] += moves into each part.  bcnt and boff hold compaction_blocks()
// ints each.
void launch_compact_moves(const u8* flag, i64 n, const i32* part, const i32* target,
                          const float* gain, i32 k, i32* bcnt, i32* boff, i64* total,
                          i32* mv_v, i32* mv_src, i32* mv_dst, float* mv_gain, i32* outc,
                          i32* inc, hipStream_t s);
inline i64 compaction_blocks() { return 4096; }
// avail[q] per part (gpu_part.cu:358-369); also zeroes *zero1 when given (the
// applied-move count that launch_commit appends to); with inc, *fit (set to 1
// beforehand) becomes 0 when some inc[q] > avail[q]
void launch_capacity(const i32* pc, const i32* outc, i32 k, i64 n, int phase, i64* avail,
                     i32* zero1, const i32* inc, i32* fit, hipStream_t s);
// every one of the M moves selected in buffer order: sk_idx[i] = i, sel[i] = 1,
// *S = M, *pads = 0
void launch_select_all(i64 M, i32* sk_idx, u8* sel, i32* S, i32* pads, hipStream_t s);
// first_done: the first tile pass already ran (launch_sort_first_dev)
void launch_sort_moves(const float* mv_gain, i64 M, i64 len, float* sk_gain, i32* sk_idx,
                       hipStream_t s, bool first_done = false);
// The sort's first tile pass with M read from the device (min(*Mp, n); the
// network length min(2^ceil(log2 M), n) as the refinement pads it), for a
// launch ahead of the host's decision; nothing when M < 2.
void launch_sort_first_dev(const float* mv_gain, const i64* Mp, i64 n, float* sk_gain,
                           i32* sk_idx, hipStream_t s);
i64 select_scratch_elems(i64 M, i32 k);
void launch_select(const i32* sk_idx, const i32* mv_dst, i64 M, i32 k, const i64* avail,
                   i32* sdst, i32* H, u8* sel, i32* S, i32* pads, hipStream_t s);
// part8 (optional) is kept equal to part's byte copy; app (optional) receives
// the ids of the applied moves, *napp their number (*napp must start at 0)
void launch_commit(const i32* sk_idx, const u8* sel, i64 M, const i32* mv_v, const i32* mv_src,
                   const i32* mv_dst, i32* part, u8* part8, i32* pc, i32 k, i32* app,
                   i32* napp, hipStream_t s);
// Accumulators for block sums: ACC_SLOTS slots, one per 128-byte line, all
// zero between uses.  Kernels add into them; launch_acc_finish writes the
// total to *out and clears the slots.
constexpr int ACC_SLOTS = 256;
constexpr int ACC_STRIDE = 16;
constexpr i64 ACC_ELEMS = (i64)ACC_SLOTS * ACC_STRIDE;
void launch_acc_finish(u64* acc, u64* out, hipStream_t s);
// ((i32*)out)[0] = *m, [1] = *fit, out[1] = *cut, ((i32*)out)[4..5] = ctr[0..1],
// ((i32*)out)[6..6+k) = pc; a null input leaves zeros (out holds 3 + (k + 1) / 2
// words)
void launch_pass_status(const i64* m, const u64* cut, const i32* ctr, const i32* pc, i32 k,
                        const i32* fit, i64* out, hipStream_t s);
// asrc[i] = source vertex of arc i (built once per level)
void launch_arc_source(DGraph g, i32* asrc, hipStream_t s);
// Edge cut over arcs [0, e_end) with e_end = max(xadj[n-1], nedges): exactly
// the arcs of the reference's row ranges (gpu_part.cu:404-417).
// (acc: accumulator slots).  part8 (optional, k <= PART8_MAXK): the byte
// copy of part, read instead of part.
void launch_cut_arcs(DGraph g, i64 e_end, const i32* asrc, const i32* part, const u8* part8,
                     i32 k, int unit, u64* acc, hipStream_t s);

// ---- balance stage (tgp_balance.hip): finest level, cap per part --------------
struct BalScratch {
  i32* bcnt;       // 4096 block counts
  i64* total;      // 1
  u64* keys;       // n
  u64* keys2;      // n
  i32* vals;       // n
  i32* vals2;      // n
  i64* start;      // k
  void* sort_tmp;  // sort_bytes
  i64 sort_bytes;
};
i64 balance_sort_bytes(i64 max_items);
// T[v * k + q] = weight of v's arcs into part q (self loops excluded)
void launch_bal_tally(DGraph g, const i32* part, i32 k, i32* T, hipStream_t s);
// best destination with room (room[q] = cap - count): mode 0 vertices of parts
// over cap (room < 0), any gain; mode 1 every vertex, positive gain
void launch_bal_best(i64 n, const i32* part, const i32* T, i32 k, const i32* room, int mode,
                     u8* cand, i32* target, i32* gain, hipStream_t s);
// winners: candidates with no neighbouring candidate of larger gain, or equal
// gain and smaller id
void launch_bal_token(DGraph g, const u8* cand, const i32* gain, u8* win, hipStream_t s);
// keep[v] = flagged and ranked (gain desc, v asc) below limit within keypart
// (synchronises the stream once to read the candidate count)
void launch_bal_rank_keep(const u8* flag, i64 n, const i32* keypart, const i32* gain, i32 k,
                          const i32* limit, u8* keep, const BalScratch& S, hipStream_t s);
// before applying a round's moves on a level with equal reverse weights: the
// tallies of the moved vertices' neighbours, moved from part[v] to target[v]
void launch_bal_tally_update(DGraph g, const u8* keep, const i32* part, const i32* target, i32 k,
                             i32* T, hipStream_t s);
// avail[q] = max(cap - pc[q], 0)
void launch_bal_avail(const i32* pc, i32 k, i64 cap, i64* avail, hipStream_t s);
// part[v] = target[v] for kept v, counts updated, *moved += number applied
void launch_bal_apply(const u8* keep, i64 n, const i32* target, i32* part, i32* pc, u64* moved,
                      hipStream_t s);

}  // namespace tgp
