/*
 * tgp_oracle.h -- CPU oracle for the multilevel k-way partitioning path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C++ restatement of the
 * reference program src/gpu_part.cu (iamrupamroyy/TheGraphPartitioning)
 * under one fixed legal thread schedule ("SEQ": every kernel launch behaves
 * as if its threads ran to completion one at a time in ascending global
 * thread index; host code runs as written).  Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py may load it, and only as the checker.
 * The product library (thegraphpartitioning_amd) never links or calls it.
 *
 * Each function cites the reference lines it restates.
 */
#ifndef TGP_ORACLE_H
#define TGP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* initial-partition modes for the coarsest graph (gpu_part.cu:803-864, 1157-1185) */
enum {
  TGPO_INIT_BLOCK = 0,   /* part[i] = i*k/n_c: deterministic stand-in for gpmetis */
  TGPO_INIT_ARRAY = 1,   /* caller-supplied values (e.g. a fixture part file)      */
  TGPO_INIT_GPMETIS = 2  /* write ./output.graph, run `gpmetis -ptype=rb`, read part file */
};

typedef struct tgpo_options {
  int32_t k;              /* number of parts                                    */
  int32_t wide;           /* 0: reference int32 limits; 1: 64-bit cut and n*k   */
  int64_t coarsen_floor;  /* 0: max(500, n/500) (gpu_part.cu:897); else override */
  int32_t init_mode;      /* TGPO_INIT_*                                        */
  int32_t repair;         /* 0: reference semantics; 1: repaired contraction
                             and stall stop (20*n_c > 19*n ends coarsening)  */
  const int32_t* init_part; /* TGPO_INIT_ARRAY values                           */
  int64_t init_part_len;
} tgpo_options;

/* trace record kinds */
enum { TGPO_TR_LEVEL = 1, TGPO_TR_INIT = 2, TGPO_TR_PASS = 3, TGPO_TR_REFINE = 4, TGPO_TR_FINAL = 5 };

typedef struct tgpo_trace {
  int32_t kind;
  int32_t level;   /* graph level (0 = input)                       */
  int32_t phase;   /* 0 refine phase, 1 balance phase               */
  int32_t pass;
  int64_t a, b, c, d;  /* LEVEL: n, m, n_c, -        INIT: n_c, npairs, cut0, -
                          PASS: M, S, cut, pads-in-prefix
                          REFINE: n, nedges, cut, passes  FINAL: n, m, cut, - */
  uint64_t hash;   /* FNV-1a over the partition after the step (PASS/REFINE/FINAL) */
} tgpo_trace;

/* error codes */
enum {
  TGPO_OK = 0,
  TGPO_EINVAL = -1,        /* bad arguments / malformed CSR                     */
  TGPO_ELIMIT = -2,        /* int32-mode reference limit exceeded (B11)         */
  TGPO_EEDGES = -3,        /* "Invalid new_coarsed_edge" (gpu_part.cu:1122)     */
  TGPO_EINITPART = -4,     /* initial partitioner failed                        */
  TGPO_EPARTFILE = -5      /* cannot open the part file (gpu_part.cu:1165)      */
};

/* ---- whole path (gpu_part.cu:878-1322) ------------------------------------ */
int tgpo_partition(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                   const int32_t* adjwgt, const tgpo_options* opt,
                   int32_t* part /* n */, int32_t* part_count /* k */,
                   int64_t* edge_cut, tgpo_trace* trace, int64_t trace_cap,
                   int64_t* trace_len);

/* ---- level functions, exposed for known-answer tests ---------------------- */
/* gpu_part.cu:55-106, 245-250 (SEQ): lexicographically-first heavy-edge matching */
int tgpo_match(int64_t n, const int64_t* xadj, const int32_t* adjncy,
               const int32_t* adjwgt, int32_t* match);
/* gpu_part.cu:109-142: roots -> coarse ids in ascending order; returns n_c */
int64_t tgpo_relabel(int64_t n, const int32_t* match, int32_t* new_ids,
                     int32_t* trace_back /* 2n capacity */);
/* gpu_part.cu:145-242: contraction with the mod-64 bucket rule; returns m_c.
   out_xadj has n_c+1 entries, out_adj/out_w need 2*xadj[n] capacity.        */
int64_t tgpo_contract(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                      const int32_t* adjwgt, const int32_t* match,
                      const int32_t* new_ids, int64_t n_c, int64_t* out_xadj,
                      int32_t* out_adj, int32_t* out_w);
/* Opt-in repaired contraction (no reference counterpart): coarse row c is
   built from its own fine vertices tb[2c], tb[2c+1], duplicates merged at
   their first occurrence, self loops dropped.                              */
int64_t tgpo_contract_repaired(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                               const int32_t* adjwgt, const int32_t* trace_back,
                               const int32_t* new_ids, int64_t n_c, int64_t* out_xadj,
                               int32_t* out_adj, int32_t* out_w);
/* gpu_part.cu:803-853: METIS text of a CSR; returns byte length (writes at
   most cap bytes) and the unique undirected pair count.                     */
int64_t tgpo_metis_text(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                        const int32_t* adjwgt, char* buf, int64_t cap,
                        int64_t* npairs);
/* gpu_part.cu:400-469 */
int64_t tgpo_edge_cut(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                      const int32_t* adjwgt, int64_t nedges, const int32_t* part,
                      int32_t k, int32_t wide);
/* gpu_part.cu:258-305: one gain pass (phase 0/1) -> gain[n], target[n] */
int tgpo_gains(int64_t n, const int64_t* xadj, const int32_t* adjncy,
               const int32_t* adjwgt, int64_t nedges, const int32_t* part,
               const int32_t* part_count, int32_t k, int32_t phase,
               float* gain, int32_t* target);
/* gpu_part.cu:472-506: the bitonic network on (gain, id) pairs of length len */
int tgpo_bitonic(float* gain, int32_t* id, int64_t len);
/* gpu_part.cu:509-759: refinement of one level; part/part_count/cut in-out  */
int tgpo_refine(int64_t n, const int64_t* xadj, const int32_t* adjncy,
                const int32_t* adjwgt, int64_t nedges, int32_t k, int32_t wide,
                int32_t* part, int32_t* part_count, int64_t* cut,
                tgpo_trace* trace, int64_t trace_cap, int64_t* trace_len);
/* gpu_part.cu:1296-1321: bytes of finalpartition.txt */
int64_t tgpo_final_text(int64_t n, const int32_t* part, int32_t k, char* buf,
                        int64_t cap);
/* Plants one deliberate mistake in the restatement so tests can show that
   the known-answer tests catch it: 0 none (the default), 1 the phase-0
   capacity clamp uses slack instead of slack/2 (gpu_part.cu:368), 2 moves are
   ranked in buffer order instead of after the network (gpu_part.cu:596-600),
   3 the repaired contraction keeps the first weight instead of the sum.     */
void tgpo_plant_mistake(int32_t which);
/* Threads of the oracle's order-independent loops (gains, candidates, token
   passing, the network's stages, the cut); results do not depend on it.
   threads <= 0 leaves the setting; returns the number now in use.          */
int32_t tgpo_set_threads(int32_t threads);
/* FNV-1a 64 over int32 values, the hash used in trace records */
uint64_t tgpo_hash(const int32_t* v, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
