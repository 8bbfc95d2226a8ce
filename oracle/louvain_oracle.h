/*
 * louvain_oracle.h -- CPU restatement of the reference's GPU Louvain method
 * (FedeFonto/louvain-fast-move-cuda), in plain C11, single thread.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library
 * and the timed CPU baseline; the product path never links, loads or calls
 * it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may use it.
 *
 * Each function cites the reference file:line whose semantics it restates.
 * Two arithmetic modes exist:
 *   LVO_GAIN_EXACT  -- integer gains 2m*(w_vc - self) + k*(tot_own - k - tot_c)
 *                      in int64 (order-identical to the true delta-Q);
 *   LVO_GAIN_REF_F32 -- the reference's expression: f32 weight difference,
 *                      f64 arithmetic, result rounded to f32
 *                      (operatorsDelta.h:19-20, :38-39).
 */
#ifndef LOUVAIN_ORACLE_H
#define LOUVAIN_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LVO_OK = 0, LVO_ERR_ARG = 1, LVO_ERR_OOM = 2, LVO_ERR_PARSE = 3,
       LVO_ERR_IO = 4, LVO_ERR_NOT_DENSE = 5, LVO_ERR_RANGE = 6 };

/* enum MODE { HASH, SORT, ADAPTIVE_SPEED, ADAPTIVE_MEMORY } (constants.h:13) */
enum { LVO_HASH = 0, LVO_SORT = 1, LVO_ADAPTIVE_SPEED = 2, LVO_ADAPTIVE_MEMORY = 3 };
/* sweep kinds: phaseOptimization.cuh:129-144 */
enum { LVO_SWEEP_FAST = 0, LVO_SWEEP_SORT = 1, LVO_SWEEP_HASH = 2 };
enum { LVO_GAIN_EXACT = 0, LVO_GAIN_REF_F32 = 1 };
/* LVO_TIE_LOWEST: lowest community id wins every tie (the deterministic rule);
 * LVO_TIE_REFERENCE: fast/sort sweeps lowest id (operators.h:30-40 with
 * Thrust's left-to-right reduce_by_key), hash sweeps highest id
 * (phaseOptimizationHash.cu:29-31 packs the id into the low bits of an
 * atomicMax key). */
enum { LVO_TIE_LOWEST = 0, LVO_TIE_REFERENCE = 1 };

typedef struct lvo_graph {
    uint32_t n;            /* vertices at this level */
    uint64_t E;            /* directed adjacency entries */
    uint64_t *row;         /* n+1 CSR offsets, entries sorted by (src,dst) */
    uint32_t *col;         /* E destinations */
    uint32_t *w;           /* E integer weights */
    uint64_t *k;           /* n weighted degrees (tot_weight_per_nodes) */
    uint64_t m2;           /* sum of k = 2m */
    double total_weight;   /* graph.total_weight: m at level 0, never updated */
    uint64_t n_links;      /* graph.n_links (undirected edge count / E'/2) */
    int64_t min_id;        /* GraphHost::min_id */
} lvo_graph;

typedef struct lvo_params {
    int mode;                 /* LVO_HASH ... */
    int gain;                 /* LVO_GAIN_* */
    int tie;                  /* LVO_TIE_* */
    int chunk_emulation;      /* 1: reproduce STEP_ROUND chunk bounds and the
                                 empty-chunk early return of the fast/sort sweeps */
    uint64_t chunk_entries;   /* STEP_ROUND (constants.h:18) */
    double sweep_threshold;   /* phaseOptimization.cuh:179 */
    double level_threshold;   /* modularityAlgorithms.h:52 */
    int64_t max_sweeps;       /* EARLY_STOP_LIMIT (constants.h:16) */
    int64_t max_total_sweeps; /* cap over all levels (bounded CPU samples); <=0: none */
    int32_t max_levels;       /* <=0: none */
    int32_t refine;           /* 1: after the level loop, project the partition onto the
                                 level-0 graph and sweep again (no fast sweep) until the
                                 sweep threshold; the reference's RefinePhase is empty
                                 (phaseRefine.cuh:1-3), so this is off by default */
} lvo_params;

typedef struct lvo_sweep_rec {
    int32_t level, sweep, kind, _pad;
    uint32_t n, n_active, n_moved, n_considered;
    uint64_t E, e_active;       /* e_active: selected entries (active src, src!=dst) */
    int64_t sigma_in;           /* sum of w over entries with comm[src]==comm[dst] */
    uint64_t sum_tot_sq;        /* sum_c tot_c^2, exact */
    double q;
    uint64_t comm_hash;         /* FNV-1a 64 over comm[] (u32 little endian) */
} lvo_sweep_rec;

typedef struct lvo_level_rec {
    int32_t level, sweeps;
    uint32_t n, n_comm;
    uint64_t E;
    double q_start, q_end;
    uint64_t best_hash;         /* FNV-1a 64 over best_communities after the level */
} lvo_level_rec;

typedef struct lvo_result {
    double modularity;
    uint32_t n_of_best_communities;
    int32_t levels, total_sweeps, _pad;
    /* caller-owned buffers (may be NULL) */
    uint32_t *best_communities;  /* n at level 0 */
    lvo_sweep_rec *sweeps;  int32_t sweeps_cap;
    lvo_level_rec *level_recs; int32_t levels_cap;
} lvo_result;

void lvo_default_params(lvo_params *p);

/* GraphHost::GraphHost (graph.cu:17-87) + GraphDevice(GraphHost&) (graph.cuh:46-99):
 * normalise each pair to (min,max), drop duplicates (first weight wins when
 * weights are given), emit both directions (a self-loop gives two (a,a)
 * entries), sort by (src,dst), subtract min_id, degree = sum of weights.
 * Returns LVO_ERR_NOT_DENSE unless ids cover [min_id, min_id+n). */
int lvo_graph_from_edges(uint64_t n_edges, const int64_t *a, const int64_t *b,
                         const uint32_t *w, lvo_graph **out);
int lvo_graph_load_edge(const char *path, int weighted, int skip_line, lvo_graph **out);
/* Build directly from a CSR (copied).  k and m2 are derived; total_weight = m2/2
 * unless total_weight > 0 is given. */
int lvo_graph_from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                       const uint32_t *w, double total_weight, lvo_graph **out);
void lvo_graph_free(lvo_graph *g);

/* Community::compute_modularity (community.h:68-106), exact integer parts. */
double lvo_modularity(const lvo_graph *g, const uint32_t *comm,
                      int64_t *sigma_in, uint64_t *sum_tot_sq);

/* One sweep's best-move decision from a partition snapshot
 * (phaseOptimizationSort.cu:28-158 / :161-379, phaseOptimizationHash.cu:38-162,
 * and the move rule of phaseOptimizationUpdate.cu:14-115).
 * tot[] must equal sum of k over members.  target[v] receives the community v
 * moves to, or comm[v] when v stays.  Returns the number of movers. */
int64_t lvo_best_moves(const lvo_graph *g, const uint32_t *comm, const int64_t *tot,
                       const uint8_t *active, int kind, const lvo_params *p,
                       uint32_t *target, uint64_t *e_active, uint32_t *n_considered);

/* The reference-rounding delta-Q of one candidate (DeltaModularityHash /
 * DeltaModularitySort, operatorsDelta.h:19-20, :38-39), as lvo_best_moves
 * evaluates it in LVO_GAIN_REF_F32 mode: w = w(v,c), self = w(v,own),
 * k = k_v, tw = total_weight.  Exported for the tests that pin it. */
float lvo_dq_ref_f32(uint64_t w, uint64_t self, uint64_t k, int64_t tot_own, int64_t tot_c,
                     double tw);

/* AggregationPhase::run (phaseAggregation.cuh:19-175): renumber non-empty
 * communities in id order (map = inclusive scan of tot>0), contract.
 * new_id[c] = map[c]-1 for non-empty c, 0xFFFFFFFF otherwise. */
int lvo_aggregate(const lvo_graph *g, const uint32_t *comm, uint32_t *new_id,
                  uint32_t *n_new, lvo_graph **out);

/* ModularityAlgorithms::Louvain (modularityAlgorithms.h:8-81). */
int lvo_louvain(const lvo_graph *g, const lvo_params *p, lvo_result *r);

/* Accessors for ctypes. */
uint32_t lvo_graph_n(const lvo_graph *g);
uint64_t lvo_graph_E(const lvo_graph *g);
double lvo_graph_total_weight(const lvo_graph *g);
int64_t lvo_graph_min_id(const lvo_graph *g);
void lvo_graph_copy(const lvo_graph *g, uint64_t *row, uint32_t *col, uint32_t *w, uint64_t *k);
uint64_t lvo_fnv1a_u32(const uint32_t *a, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
