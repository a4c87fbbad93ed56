/*
 * tgp.h -- C ABI of the MI355X multilevel k-way graph partitioner.
 *
 * The reference (iamrupamroyy/TheGraphPartitioning, src/gpu_part.cu) exposes
 * no library API: its only interface is the process contract
 *     gpu_part <dir> <nvtxs> <nedges> <k>     (gpu_part.cu:878-888)
 * reading <dir>/row.txt, column.txt, weight.txt (gpu_part.cu:905-908, 943-945)
 * and writing ./output.graph, ./output.graph.part.<k> (via gpmetis) and
 * ./finalpartition.txt (gpu_part.cu:829-858, 1163, 1296-1321).
 * The `gpu_part` CLI built from this library keeps that contract; the entry
 * points below are what the CLI, the tests and any FFI caller bind.  Each one
 * names the part of gpu_part.cu it replaces.
 *
 * Plain C types only: pointers and sizes, no torch or C++ types.  The caller
 * owns every buffer passed in.  Errors are negative return codes; the library
 * never calls exit().
 */
#ifndef TGP_H
#define TGP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------------- */
enum {
  TGP_OK = 0,
  TGP_EINVAL = -1,      /* bad argument or malformed CSR (offsets not monotone,
                           columns outside [0,n), xadj[n] != nedges)            */
  TGP_ELIMIT = -2,      /* a reference int32 limit is exceeded in int32 mode:
                           n*k >= 2^31 (gpu_part.cu:531), n or m > INT_MAX     */
  TGP_EEDGES = -3,      /* "Invalid new_coarsed_edge" (gpu_part.cu:1121-1125)  */
  TGP_EINITPART = -4,   /* initial partitioner failed (gpu_part.cu:857-863)    */
  TGP_EPARTFILE = -5,   /* part file missing (gpu_part.cu:1163-1168)           */
  TGP_EHIP = -6,        /* HIP runtime error (replaces CUDA_CHECK, gpu_part.cu:40-47) */
  TGP_ENOMEM = -7,      /* device or host allocation failed                    */
  TGP_EIO = -8,         /* file I/O error (gpu_part.cu:786-800, 830-833, 1306) */
  TGP_EDIFF = -9        /* tgp_batch_autotune: two engine settings gave
                           different results (no counterpart in the reference) */
};

/* ---- graphs ---------------------------------------------------------------- */
/* CSR as the reference reads it (row.txt / column.txt / weight.txt,
 * gpu_part.cu:786-800, 943-955), with 64-bit offsets so one type serves both
 * the reference's int32 limits and graphs past 2^31 arcs.                    */
typedef struct tgp_graph {
  int64_t n;              /* vertices                                           */
  int64_t m;              /* arcs (the reference's nedges); xadj[n] == m       */
  const int64_t* xadj;    /* n+1 row offsets                                   */
  const int32_t* adjncy;  /* m column ids in [0,n)                             */
  const int32_t* adjwgt;  /* m edge weights                                    */
} tgp_graph;

/* Initial partitioner of the coarsest graph.  Replaces the fixed seam
 * convertCSRtoMETISandRun + part-file read (gpu_part.cu:803-864, 1157-1174).
 * `coarsest` is host memory; `npairs` is the unique undirected pair count the
 * reference writes in the METIS header; `part` (coarsest->n entries) is
 * written by the callee.  Out-of-range values are mapped to 0 afterwards, as
 * gpu_part.cu:1172 does.  Return 0 on success.                               */
typedef int (*tgp_initpart_fn)(const tgp_graph* coarsest, int64_t npairs, int32_t k,
                               int32_t* part, void* user);

/* The reference seam itself: writes ./output.graph in the current directory
 * (byte-identical to gpu_part.cu:829-853), runs
 * `gpmetis -ptype=rb ./output.graph <k>` through /bin/sh and reads
 * ./output.graph.part.<k> with the reference's rules (gpu_part.cu:1163-1174).
 * `user` may point to an int that receives system()'s status.                */
int tgp_initpart_gpmetis(const tgp_graph* coarsest, int64_t npairs, int32_t k,
                         int32_t* part, void* user);
/* The same seam in a working directory of its own, so that several graphs
 * can call gpmetis at once (the reference's fixed names, gpu_part.cu:829,
 * 1163, allow one per directory): `user` points to a tgp_gpmetis_dir; writes
 * <dir>/output.graph (the same bytes), runs
 * `gpmetis -ptype=rb '<dir>/output.graph' <k>` and reads
 * <dir>/output.graph.part.<k> with the same rules.  The directory must exist
 * and its name may not contain a single quote.  No counterpart in the
 * reference.                                                                 */
typedef struct tgp_gpmetis_dir {
  const char* dir;
  int32_t status;          /* out: system()'s status                           */
  int32_t reserved;
} tgp_gpmetis_dir;
int tgp_initpart_gpmetis_dir(const tgp_graph* coarsest, int64_t npairs, int32_t k,
                             int32_t* part, void* user);
/* Deterministic stand-in for hosts without METIS: part[i] = i*k/n_c.  No
 * counterpart in the reference (which always calls gpmetis).                 */
int tgp_initpart_block(const tgp_graph* coarsest, int64_t npairs, int32_t k,
                       int32_t* part, void* user);

typedef struct tgp_options {
  int32_t k;               /* parts (the reference's num_partitions)           */
  int32_t wide;            /* 0: reference int32 semantics (cut wraps in int32,
                              n*k < 2^31); 1: 64-bit cut, no n*k limit          */
  int64_t coarsen_floor;   /* 0: max(500, n/500) (gpu_part.cu:897); else the
                              vertex count below which coarsening stops         */
  tgp_initpart_fn initpart;/* NULL: tgp_initpart_gpmetis (the reference seam)  */
  void* initpart_user;
  int32_t device_input;    /* 1: xadj/adjncy/adjwgt are device pointers (checked
                              on the device as a host CSR is on the host:
                              TGP_EINVAL when malformed)                        */
  int32_t profile;         /* 1: time kernels with HIP events into tgp_stats    */
  int32_t check_level;     /* 1: after every pass copy the partition to host and
                              hash it into the trace (slow; tests only)         */
  int32_t repair;          /* 0: the reference's semantics, bugs included.
                              1: opt-in repaired coarsening (no counterpart in
                              the reference): coarse row c is built from its own
                              fine vertices with all duplicates merged, and
                              coarsening stops once a level keeps more than 95%
                              of its vertices (20*n_c > 19*n).  Refinement is
                              unchanged.                                        */
  /* Optional refinement trace (no counterpart in the reference): called on the
   * host with the partition of a level at the start of each phase (pass = -1)
   * and after each committed pass; used to record the training data of the
   * learned move proposer (tgp_learn.h).  Slow: a copy per pass.            */
  void (*pass_dump)(int32_t level, int32_t phase, int32_t pass, int64_t n,
                    const int32_t* part, void* user);
  void* pass_dump_user;
  /* Optional learned move proposer (tgp_learned_model; no counterpart in the
   * reference): when set, the finest level runs learned passes before its
   * refinement (see tgp_learned_model).                                      */
  const struct tgp_learned_model* learned;
  /* Imbalance bound of a balance stage at the finest level (no counterpart
   * in the reference, whose refinement keeps only 5 % + slack per level and
   * ends far above it on power-law graphs).  0: off, the reference's result.
   * eps > 0: after the finest level's refinement, with cap =
   * floor((1 + eps) * ceil(n / k)): balance rounds move vertices out of the
   * parts above cap (each to the part with room of largest exact integer
   * gain; ranked per destination, then per source, by gain desc and vertex
   * id asc, within room and excess) until every part holds at most cap; then
   * the learned passes, when a model is given, with capacity cap - count;
   * then refine passes of positive exact gains that keep every part within
   * cap (an independent set of winners: no neighbouring candidate with a
   * larger gain or the same gain and a smaller id; ranked per destination
   * within room).  part_count then holds the final partition's counts.
   * k <= 1024.                                                              */
  float balance_eps;
} tgp_options;

/* The learned move proposer's weights (device memory) for the GNN of
 * tgp_learn.h.  Each learned pass at the finest level: features -> GNN ->
 * for each vertex its best-scoring other part q, proposed when its logit is
 * within `margin` of the own part's and admitted only when the exact integer
 * gain T_v[q] - T_v[part_v] is positive; the reference's token rule
 * on those gains, move buffer, capacity (gpu_part.cu:316-369), sort, select
 * and commit (472-397).  Learned passes stop when fewer than min_admit * n
 * moves are admitted (or after max_passes); the level's ordinary refinement
 * follows (the fallback).  Needs k equal to the partition's k, k <= 256,
 * hidden > 0 and features >= 2k + 2 (else TGP_EINVAL).  The library reads
 * the weights on its own stream: work that writes them must have finished
 * (or be ordered before the call) when tgp_partition is called.            */
typedef struct tgp_learned_model {
  int32_t k, features, hidden, max_passes;
  const uint16_t* W1a;  /* features x hidden, BF16 */
  const uint16_t* W1b;  /* features x hidden */
  const uint16_t* W2a;  /* hidden x k */
  const uint16_t* W2b;  /* hidden x k */
  const float* b1;      /* hidden, FP32 */
  const float* b2;      /* k */
  float min_admit;
  float margin;         /* a proposal needs logit(q) > logit(part_v) - margin */
  int32_t infer_every;  /* the GNN runs on passes 0, r, 2r, ... (r = infer_every;
                           <= 1: every pass); the passes between propose from
                           the last logits, still gated by the current exact
                           gains and capacities                              */
} tgp_learned_model;

/* trace record kinds */
enum { TGP_TRACE_LEVEL = 1, TGP_TRACE_INIT = 2, TGP_TRACE_PASS = 3, TGP_TRACE_REFINE = 4,
       TGP_TRACE_FINAL = 5, TGP_TRACE_BALANCE = 6 };

/* trace records (same layout the oracle emits, used by differential tests)  */
typedef struct tgp_trace {
  int32_t kind;            /* 1 LEVEL, 2 INIT, 3 PASS, 4 REFINE, 5 FINAL        */
  int32_t level, phase, pass;
  int64_t a, b, c, d;      /* LEVEL: n, m, n_c, m_c    INIT: n_c, npairs, cut0
                              PASS: moves M, selected S, cut, pads in prefix
                              REFINE: n, nedges, cut, passes  FINAL: n, m, cut
                              BALANCE: moves in balance rounds, moves in refine
                              passes, cut, largest part (pass = rounds)        */
  uint64_t hash;           /* FNV-1a of the partition (only with check_level)  */
} tgp_trace;

#define TGP_NKERNELS 16
typedef struct tgp_stats {
  int32_t levels;                  /* coarsening levels built (reference `pass`) */
  int32_t passes;                  /* refinement passes run, all levels          */
  int64_t coarsest_n, coarsest_npairs;
  int64_t match_rounds;            /* parallel rounds of the exact matching      */
  double  seconds_total;           /* wall time of tgp_partition                 */
  double  seconds_initpart;        /* initial partitioner callback               */
  /* per-kernel totals when options.profile = 1 (see tgp_kernel_name)         */
  double  kernel_ms[TGP_NKERNELS];
  int64_t kernel_launches[TGP_NKERNELS];
  double  kernel_bytes[TGP_NKERNELS]; /* algorithmic bytes moved, per kernel    */
  tgp_trace* trace;                /* caller buffer for trace records or NULL    */
  int64_t trace_cap, trace_len;
  int64_t fail_level, fail_value;  /* TGP_EEDGES: level and the m_c it rejected  */
  int64_t sorts_skipped;           /* passes whose move sort was skipped: every
                                      move fit its destination and the network
                                      would have sorted them all, so it could not
                                      change the selection                       */
  int64_t level0_form;             /* what the symmetry check found on level 0:
                                      8 = it ran, | 1 rows strictly ascending,
                                      | 2 symmetric, | 4 every arc between
                                      distinct vertices weighs 1; 0 = not run  */
} tgp_stats;

/* Partitions `g` into o->k parts.  Replaces main()'s pipeline
 * (gpu_part.cu:960-1322): coarsening, initial partition, uncoarsening with
 * refinement, final edge cut.  Outputs (host memory, caller-owned):
 *   part[n]          final partition (finalpartition.txt content)
 *   part_count[k]    the reference's d_partCount at exit, i.e. the counts of
 *                    the last committed refinement pass (gpu_part.cu:1338-1341)
 *   edge_cut         final cut, floor(arc cut / 2) (gpu_part.cu:1285, 468)
 * Default device 0, one HIP stream.                                          */
int tgp_partition(const tgp_graph* g, const tgp_options* o, int32_t* part,
                  int32_t* part_count, int64_t* edge_cut, tgp_stats* stats);

/* ---- batched form (no counterpart in the reference, which partitions one
 * graph per process, gpu_part.cu:878-1344) ---------------------------------
 * Partitions `count` independent graphs.  Each item is partitioned exactly as
 * tgp_partition(&it->g, &it->o, it->part, it->part_count, &it->edge_cut,
 * it->stats) would do alone, so every result is bit-identical to its solo
 * run; up to `workers` items (1..64) run at once, each on its own HIP stream
 * and device pool.  Initial-partition callbacks are serialised (the gpmetis
 * seam uses fixed file names in the working directory), except
 * tgp_initpart_block and tgp_initpart_gpmetis_dir, which run concurrently
 * (give each item its own directory).  Returns TGP_OK when
 * every item succeeded, else the first failing item's code; each item's own
 * code is in it->rc.                                                         */
typedef struct tgp_batch_item {
  tgp_graph g;             /* host CSR, or device CSR when o.device_input = 1  */
  tgp_options o;
  int32_t* part;           /* out: part[g.n]                                    */
  int32_t* part_count;     /* out: part_count[o.k]                              */
  tgp_stats* stats;        /* optional                                          */
  int64_t edge_cut;        /* out                                               */
  int32_t rc;              /* out: TGP_OK or the item's error code              */
  int32_t reserved;
} tgp_batch_item;
int tgp_partition_batch(tgp_batch_item* items, int64_t count, int32_t workers);

/* Batched-engine tuning (no counterpart in the reference).  The engine's
 * refinement launch runs with a number of lanes per light vertex (2, 4 or 8)
 * and a sort tile held in LDS (2048, 4096 or 8192 entries), and each graph
 * with a heavy-row threshold (rows longer than it are refined by a whole
 * wave) chosen by its mean degree m/n.  None of them changes a result.  They
 * come from the table file TGP_BATCH_TUNE names, else batch_tune.txt in the
 * library's directory, else the built-in 4 lanes, 4096 entries, threshold
 * 96; TGP_BATCH_SETTING="lanes,heavy,tile" sets one setting for every graph.
 * Table lines ('#' starts a comment):
 *   lanes <2|4|8>
 *   sort_tile <2048|4096|8192>
 *   heavy_default <threshold>
 *   heavy <mean degree from> <mean degree below> <threshold>   (first match)
 * A malformed table makes tgp_partition_batch return TGP_EINVAL with the
 * file and line in tgp_last_error().
 *
 * tgp_batch_tuning: the setting the engine would use for a graph of this
 * mean degree (reads the table as the engine does; needs no GPU).          */
int tgp_batch_tuning(double mean_degree, int32_t* lanes, int32_t* heavy, int32_t* sort_tile);

/* tgp_batch_autotune: measures the settings on the items' graphs (a sample
 * of the workload; the items the engine takes, as in tgp_partition_batch)
 * and writes the table to table_path (if not null).  One coarsening launch,
 * then the refinement launch under every setting -- lanes {2, 4, 8} x
 * thresholds {48, 96, 192, 384, 768, 1536} for each mean-degree class [0, 12), [12, 20),
 * [20, inf), then tiles {2048, 4096, 8192} -- scored by the sum of the
 * graphs' workgroup times.  Every setting's results must equal the first's
 * (else TGP_EDIFF).  The items receive the chosen setting's results (rc
 * TGP_OK; TGP_EINVAL for graphs the engine does not take).  *report (if not
 * null) receives the measurements and the table as text; free it with
 * tgp_free.                                                                  */
int tgp_batch_autotune(tgp_batch_item* items, int64_t count, const char* table_path,
                       char** report);

/* The balance stage of tgp_options.balance_eps alone, on a partition the
 * caller holds (no counterpart in the reference): part[n] (host memory, in
 * and out; values outside [0, k) count as part 0) is balanced to at most
 * floor((1 + eps) * ceil(n / k)) vertices per part and refined within that
 * bound exactly as the last step of tgp_partition with balance_eps does;
 * part_count[k] and edge_cut receive the result's counts and cut.  Uses
 * o->k (<= 1024), o->balance_eps (> 0), o->device_input, o->wide and
 * o->check_level; the graph is g.  For callers that already hold a
 * partition (from another partitioner, or an earlier call without a bound). */
int tgp_balance(const tgp_graph* g, const tgp_options* o, int32_t* part, int32_t* part_count,
                int64_t* edge_cut, tgp_stats* stats);

/* ---- partition without allocating (no counterpart in the reference) -------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * partition as tgp_partition (bit for bit, same arguments and outputs), but
 * every device buffer it needs -- the uploaded CSR when the input is host
 * memory, the coarse levels, refinement state, tallies -- comes from one
 * static device array of TGP_NOALLOC_SCRATCH_BYTES per GPU, loaded with the
 * library and held until the process ends; it calls no hipMalloc, hipFree or
 * other allocating routine.  Graphs above TGP_NOALLOC_MAX_N vertices,
 * TGP_NOALLOC_MAX_M arcs or TGP_NOALLOC_MAX_K parts are rejected (TGP_EINVAL,
 * message in tgp_last_error); a partition that still needs more than the
 * scratch holds fails with TGP_ENOMEM.  No learned proposer or profiling.
 * The first call on a device sets up that device's context (its stream and
 * the scratch address).  Calls share the device's scratch, which keeps the
 * last call's contents: the library clears whatever must start at zero, and
 * calls on one device are serialised by a lock (calls that overlapped on one
 * GPU's scratch would race).                                                 */
#define TGP_NOALLOC_SCRATCH_BYTES ((int64_t)512 << 20)
#define TGP_NOALLOC_MAX_N ((int64_t)1 << 20)
#define TGP_NOALLOC_MAX_M ((int64_t)1 << 23)
#define TGP_NOALLOC_MAX_K 256
int tgp_partition_noalloc(const tgp_graph* g, const tgp_options* o, int32_t* part,
                          int32_t* part_count, int64_t* edge_cut, tgp_stats* stats);

const char* tgp_strerror(int code);
/* Detail of the last failure on the calling thread ("" if none).            */
const char* tgp_last_error(void);
const char* tgp_kernel_name(int id);
/* Returns cached device memory of the calling process to the HIP runtime.   */
void tgp_release_device_memory(void);
/* Copies a host CSR into device memory (so tgp_partition can run with
 * device_input = 1 on a graph already resident in HBM); no counterpart in the
 * reference, whose main() uploads inline (gpu_part.cu:950-955).             */
int tgp_graph_upload(const tgp_graph* host, tgp_graph* dev);

/* The move sort of one refinement pass, on device buffers: key/idx[0..len)
 * receive the reference's bitonic network (bitonicSortStep +
 * customSortMoves, gpu_part.cu:472-506) applied to (gain[i], i) for i < M
 * and (-1e9, -1) pads for M <= i < len (len <= 2^31), enqueued on `stream`
 * (a hipStream_t; NULL = the default stream).  The same kernels the
 * partitioner runs; exported so that the network can be checked on its own
 * at lengths the partition tests do not reach.                              */
int tgp_sort_moves(const float* gain, int64_t M, int64_t len, float* key, int32_t* idx,
                   void* stream);
void tgp_graph_device_free(tgp_graph* dev);

/* ---- files of the reference contract ------------------------------------- */
/* Reads <dir>/row.txt, column.txt, weight.txt holding exactly n+1, m, m
 * decimal ints (gpu_part.cu:786-800, 905-908, 943-945).  Arrays are malloc'd;
 * release with tgp_graph_free.                                                */
int tgp_read_graph_dir(const char* dir, int64_t n, int64_t m, tgp_graph* out);
/* Writes ./finalpartition.txt bytes (gpu_part.cu:1294-1321) to `path`.       */
int tgp_write_final_partition(const char* path, const int32_t* part, int64_t n, int32_t k);
/* METIS text of a CSR (gpu_part.cu:803-853) into a malloc'd buffer.          */
int tgp_metis_text(const tgp_graph* g, char** text, int64_t* len, int64_t* npairs);

/* ---- input and output formats -------------------------------------------- */
/* The README form's inputs (README.md:133-141): row.txt and column.txt only,
 * n = count(row) - 1, m = count(column), unit weights.                      */
int tgp_read_csr_files(const char* row_path, const char* col_path, tgp_graph* out);
/* Binary CSR file (no counterpart in the reference): "TGPCSR01", int64 n,
 * int64 m, int64 xadj[n+1], int32 adjncy[m], int32 adjwgt[m], little-endian. */
int tgp_write_bcsr(const char* path, const tgp_graph* g);
int tgp_read_bcsr(const char* path, tgp_graph* out);
/* METIS-style partition file: one line per vertex with its part (no
 * counterpart in the reference's outputs, which are finalpartition.txt).     */
int tgp_write_part_file(const char* path, const int32_t* part, int64_t n);
/* preprocess/conversion.py:30-110.  tgp_convert_edges reads an edge list by
 * extension (.mtx/.mmio Matrix Market, arcs column -> row; .tsv; .txt with 4
 * header lines; .tsv_1 with 1), drops self loops, makes every edge
 * bidirectional once, removes isolated vertices (order-preserving relabel),
 * unit weights, rows ascending (DGL's own column order is not specified).
 * tgp_write_conversion_files writes row.txt, column.txt, weight.txt,
 * output_output.csr and temp.graph into `dir` in the script's text forms.  */
int tgp_convert_edges(const char* path, tgp_graph* out);
int tgp_write_conversion_files(const tgp_graph* g, const char* dir);

/* ---- synthetic workloads (no counterpart in the reference) ---------------- */
/* weight_mode 0: unit weights; 1: 1 + hash(min,max) mod 16                   */
int tgp_gen_rmat(int32_t scale, int32_t edge_factor, uint64_t seed, int32_t weight_mode,
                 tgp_graph* out);
int tgp_gen_grid3d(int32_t nx, int32_t ny, int32_t nz, tgp_graph* out);
int tgp_gen_er(int64_t n, double mean_degree, uint64_t seed, tgp_graph* out);
/* R-MAT on n vertices (any n >= 2): samples of the scale ceil(log2 n) model
 * whose two endpoints are both below n, edge_factor * n of them.            */
int tgp_gen_rmat_n(int64_t n, int32_t edge_factor, uint64_t seed, tgp_graph* out);
void tgp_graph_free(tgp_graph* g);
void tgp_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
