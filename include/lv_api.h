/*
 * lv_api.h -- C ABI of the MI355X (gfx950) Louvain library (liblouvain_amd.so).
 *
 * The reference (FedeFonto/louvain-fast-move-cuda) exposes only header-only
 * C++ (community.h, modularityAlgorithms.h, phase*.cuh); it has no FFI.
 * Each entry point below names the reference interface it replaces.  The C++
 * mirror of those names lives in community.h / modularityAlgorithms.h /
 * phaseOptimization.h / phaseAggregation.h / phaseRefine.h next to this file.
 *
 * Conventions: every call returns an lv_status (nothing throws across the
 * ABI); output buffers are caller-owned; handles are opaque and used by one
 * host thread at a time; streams are passed as void* (a hipStream_t, NULL for
 * a library-owned stream); every HIP call is checked.
 */
#ifndef LV_API_H
#define LV_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lv_status {
    LV_OK = 0,
    LV_ERR_ARG = 1,        /* bad argument */
    LV_ERR_OOM = 2,        /* device or host allocation failed (the reference's std::bad_alloc, main.cu:26-45) */
    LV_ERR_PARSE = 3,      /* no edge parsed (the reference calls exit(0), graph.cu:52-55) */
    LV_ERR_IO = 4,         /* file cannot be opened */
    LV_ERR_NOT_DENSE = 5,  /* vertex ids do not cover [min_id, min_id+n) (graph.cuh:66-76 assumes they do) */
    LV_ERR_RANGE = 6,      /* total weight 2m >= 2^31 or more than 2^32-1 vertices */
    LV_ERR_HIP = 7,        /* a HIP runtime call failed */
    LV_ERR_NO_DEVICE = 8,  /* no GPU visible */
    LV_ERR_STATE = 9       /* call out of order */
} lv_status;

/* enum MODE { HASH, SORT, ADAPTIVE_SPEED, ADAPTIVE_MEMORY } (constants.h:13) */
typedef enum lv_mode { LV_HASH = 0, LV_SORT = 1, LV_ADAPTIVE_SPEED = 2, LV_ADAPTIVE_MEMORY = 3 } lv_mode;
/* sweep kinds scheduled by OptimizationPhase::optimize (phaseOptimization.cuh:116-155) */
typedef enum lv_sweep_kind { LV_SWEEP_FAST = 0, LV_SWEEP_SORT = 1, LV_SWEEP_HASH = 2 } lv_sweep_kind;
/* LV_GAIN_EXACT: integer gain 2m*(w_vc-self) + k*(tot_own-k-tot_c) in int64;
 * LV_GAIN_REF_F32: the reference's f32-rounded delta-Q (operatorsDelta.h:19-20) */
typedef enum lv_gain { LV_GAIN_EXACT = 0, LV_GAIN_REF_F32 = 1 } lv_gain;
/* LV_TIE_LOWEST: lowest community id wins ties in every sweep;
 * LV_TIE_REFERENCE: lowest in fast/sort sweeps, highest in hash sweeps
 * (phaseOptimizationHash.cu:29-31) */
typedef enum lv_tie { LV_TIE_LOWEST = 0, LV_TIE_REFERENCE = 1 } lv_tie;

typedef struct lv_params {
    int32_t mode;             /* lv_mode */
    int32_t gain;             /* lv_gain */
    int32_t tie;              /* lv_tie */
    int32_t chunk_emulation;  /* 1: reproduce the STEP_ROUND chunking and the fast/sort
                                 sweeps' empty-chunk early return (phaseOptimizationSort.cu:81-86, :235-240) */
    uint64_t chunk_entries;   /* STEP_ROUND (constants.h:18) */
    double sweep_threshold;   /* MODULARITY_CONVERGED_THRESHOLD for the sweep loop (phaseOptimization.cuh:179) */
    double level_threshold;   /* ... and for the level loop (modularityAlgorithms.h:52) */
    int64_t max_sweeps;       /* EARLY_STOP_LIMIT (constants.h:16) */
    int32_t max_levels;       /* <= 0: no limit */
    int32_t profile;          /* 1: time every best-move launch with HIP events;
                                 2: time the sweeps, the best-move passes and the
                                    degree 257-1024 bucket's launches only (the one
                                    that runs with the device to itself);
                                 3: as 1, with the best-move pass on one stream
                                    (each bucket timed alone; slower runs) */
    int32_t autotune;         /* 1: on a community's first fast and first hash sweep, time
                                 each degree bucket's tile candidates and keep the fastest
                                 (decisions, hence results, do not depend on the tile) */
    int32_t refine;           /* 1: after the level loop, sweep the final partition again on
                                 the level-0 graph (RefinePhase; the reference's is empty,
                                 phaseRefine.cuh:1-3, so 0 by default).  Not for batches
                                 or the server. */
    uint32_t bucket_max_deg[8]; /* degree bounds of the best-move kernels' LDS buckets 0-7
                                   (rows above bucket_max_deg[7] take the hub path):
                                   non-decreasing, each at most the default
                                   {4, 8, 16, 32, 64, 256, 1024, 4096} except bucket 5's,
                                   which may reach 512 (its tables then take 772 slots),
                                   the last at least 2 (else LV_ERR_ARG).  All zero: the default, or with
                                   autotune = 1 the fastest of three sets, timed once per
                                   community on its level-0 fast sweep.  Results do not
                                   depend on the bounds.  Not for the server (default). */
} lv_params;

typedef struct lv_sweep_rec {
    /* pre_decided: 1 when the sweep's best moves were decided by the
     * contraction that built the level (the first, fast sweep of a level
     * above 0 in lv_community_run): no best-move pass ran, move_ms and
     * move_bytes are 0 and that work is in the contraction's time */
    int32_t level, sweep, kind, pre_decided;
    uint32_t n, n_active, n_moved, n_considered;
    uint64_t E, e_active;      /* e_active: entries of active vertices with src != dst */
    int64_t sigma_in;          /* weight of entries inside communities */
    uint64_t sum_tot_sq;       /* sum over communities of tot_c^2 */
    double q;
    uint64_t comm_hash;        /* FNV-1a 64 over comm[] (0 unless requested) */
    /* with lv_params.profile set (else 0): HIP-event time of the sweep's
     * best-move pass and of the whole sweep, the distinct (vertex, neighbour
     * community) candidates it scored, and its algorithmic bytes
     * beta*e_active + 8*candidates + 28*n_active + n (beta = 8 at unit
     * weights, 12 weighted) */
    double move_ms, sweep_ms;
    uint64_t candidates, move_bytes;
    /* mover_entries (always): entries of the rows of the vertices that moved.
     * With profile set: HIP-event time of the commit (apply, marking, the
     * change of sigma_in, next lists: sweep_ms - move_ms) and its algorithmic
     * bytes 8*n_active + 24*n_moved + 5*mover_entries + n (apply + push-form
     * marking). */
    double commit_ms;
    uint64_t commit_bytes, mover_entries;
} lv_sweep_rec;

typedef struct lv_level_rec {
    int32_t level, sweeps;
    uint32_t n, n_comm;
    uint64_t E;
    double q_start, q_end;
    uint64_t best_hash;        /* 0 unless requested */
    /* with profile set and a contraction after the level (else 0): HIP-event
     * time from renumbering to the contracted graph (host planning steps
     * included), the contracted graph's size, and the algorithmic bytes
     * beta*E + 8*E_new + 12*n + 12*n_comm (beta = 8 at unit weights, 12
     * weighted: col, w, new id of the destination; the contracted entries;
     * per vertex its row offset and new id, per community its offset and
     * degree) */
    double contract_ms;
    uint64_t E_new, contract_bytes;
} lv_level_rec;

typedef struct lv_result {
    double modularity;
    uint32_t n_of_best_communities;
    int32_t levels, total_sweeps, hash_trace;  /* hash_trace (in): 1 = fill comm_hash/best_hash */
    uint32_t *best_communities;  /* caller buffer of n (level-0 vertices), may be NULL */
    lv_sweep_rec *sweeps; int32_t sweeps_cap;
    lv_level_rec *level_recs; int32_t levels_cap;
    double seconds;              /* device-resident CSR to final partition, all levels */
    uint64_t traversed_entries;  /* sum of e_active over all sweeps */
} lv_result;

typedef struct lv_kernel_stats {
    /* best-move passes (all degree buckets), summed over the sweeps timed
     * when lv_params.profile is 1 or 2 */
    uint64_t move_launches;
    double move_ms;             /* HIP-event time of the passes on the library stream */
    uint64_t move_bytes;        /* algorithmic bytes of those launches */
    uint64_t move_entries;      /* entries traversed by those launches */
    double sweep_ms;            /* whole sweeps (move + commit + modularity), event timed */
    uint64_t sweeps;
    /* per best-move kernel variant v = ((bucket*2 + fast)*2 + weighted)*3 + tile,
     * bucket 0-7, tile = index of the bucket's tile candidate (one kernel launch
     * each; lv_kernel_variant_name gives the kernel name).  Buckets 5 and 7
     * share the device with others while they run, so their times include
     * that; with profile = 2 only bucket 6's variants are recorded. */
    uint64_t variant_launches[96];
    double variant_ms[96];
    uint64_t variant_bytes[96];
    uint64_t variant_entries[96];
    /* per degree bucket b (0-7, 8 = hub rows) and sweep kind f (0 hash/sort,
     * 1 fast), index 2*b + f: launches (one per sweep), event time, bytes and
     * entries.  With profile = 3 the pass runs without side streams, so each
     * bucket is timed with the device to itself. */
    uint64_t bucket_launches[18];
    double bucket_ms[18];
    uint64_t bucket_bytes[18];
    uint64_t bucket_entries[18];
    /* commit passes (lv_sweep_rec.commit_*) and contractions
     * (lv_level_rec.contract_*) summed over the timed sweeps and levels;
     * commit_entries = mover entries, contract_entries = entries contracted */
    uint64_t commit_passes;
    double commit_ms;
    uint64_t commit_bytes, commit_entries;
    uint64_t contractions;
    double contract_ms;
    uint64_t contract_bytes, contract_entries;
} lv_kernel_stats;

typedef struct lv_graph lv_graph;          /* GraphHost (graphHost.cuh:7-18): host CSR */
typedef struct lv_community lv_community;  /* Community (community.h:15-107): device state */

void lv_default_params(lv_params *p);
const char *lv_status_string(lv_status s);
lv_status lv_device_count(int *count);

/* GraphHost::GraphHost(std::string name, bool weighted, int skip_line) (graph.cu:17-87).
 * weighted = 1 reads a third integer column (an extension: the reference
 * leaves weighted parsing as a TODO, graph.cu:32-34). */
lv_status lv_graph_load_edge(const char *path, int weighted, int skip_line, lv_graph **out);
/* Same semantics from in-memory pairs (w may be NULL for unit weights). */
lv_status lv_graph_from_edges(uint64_t n_edges, const int64_t *a, const int64_t *b,
                              const uint32_t *w, lv_graph **out);
/* Direct CSR (row[n+1], col[E], w[E] or NULL); total_weight <= 0 means sum(w)/2,
 * any other value must equal sum(w)/2 (else LV_ERR_ARG). */
lv_status lv_graph_from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                            const uint32_t *w, double total_weight, lv_graph **out);
void lv_graph_free(lv_graph *g);

/* ---- .edge ingest on the GPU (graph.cu:17-87 built on the device) ----
 * The same graph lv_graph_load_edge builds (same reads, dedupe, min_id shift,
 * dense-id check and errors), parsed and assembled in HBM: the file is read
 * in pinned chunks overlapped with their upload, tokenised, normalised,
 * radix sorted and compacted into the CSR on the device. */
typedef struct lv_ingest_info {
    int64_t min_id;
    uint32_t n;
    uint64_t E;
    double total_weight;
    double seconds_read_upload;   /* file read + copy to the device (overlapped) */
    double seconds_tokenize;
    double seconds_build;
} lv_ingest_info;
/* Straight into a community (no host CSR); info may be NULL. */
lv_status lv_community_load_edge(const char *path, int weighted, int skip_line, void *stream,
                                 lv_community **out, lv_ingest_info *info);
/* Into a host graph handle (the device CSR is copied back); info may be NULL. */
lv_status lv_graph_load_edge_gpu(const char *path, int weighted, int skip_line, lv_graph **out,
                                 lv_ingest_info *info);
/* Writes edges as .edge text, "a b" or "a b w" per line (host, multithreaded). */
lv_status lv_write_edge(const char *path, const int64_t *a, const int64_t *b, const uint32_t *w,
                        uint64_t m);
lv_status lv_graph_info(const lv_graph *g, uint32_t *n, uint64_t *E, double *total_weight,
                        int64_t *min_id);
lv_status lv_graph_copy_csr(const lv_graph *g, uint64_t *row, uint32_t *col, uint32_t *w);

/* Community(GraphHost&) (community.h:31-46): uploads the CSR (GraphDevice,
 * graph.cuh:46-99), identity partition, totals = degrees, modularity. */
lv_status lv_community_create(const lv_graph *g, void *stream, lv_community **out);
void lv_community_destroy(lv_community *c);
/* OptimizationPhase::run(Community&, MODE) (phaseOptimization.cuh:184-206) */
lv_status lv_optimization_phase(lv_community *c, const lv_params *p);
/* AggregationPhase::run(Community&, bool aggregate, MODE) (phaseAggregation.cuh:19-175) */
lv_status lv_aggregation_phase(lv_community *c, int aggregate, const lv_params *p);
/* RefinePhase::run() (phaseRefine.cuh:1-3): a no-op, as in the reference */
lv_status lv_refine_phase(lv_community *c);
/* The opt-in refinement on a community whose level loop has ended: the final
 * partition swept again on the level-0 graph, best_communities renumbered. */
lv_status lv_refine(lv_community *c, const lv_params *p, lv_result *r);
/* Community::modularity (community.h:26) */
double lv_modularity(const lv_community *c);
/* Community::best_communities / n_of_best_communities (community.h:20,25) */
lv_status lv_best_communities(const lv_community *c, uint32_t *out, size_t n, uint32_t *n_comm);
/* ModularityAlgorithms::Louvain(GraphHost&, MODE) (modularityAlgorithms.h:8-81) */
lv_status lv_louvain(const lv_graph *g, const lv_params *p, lv_result *r);

/* Device-resident re-runs (no reference counterpart; the reference builds a
 * new Community per call, modularityAlgorithms.h:18): reset returns to the
 * resident level-0 CSR with the identity partition; run executes the level
 * loop of ModularityAlgorithms::Louvain on the community as it stands. */
lv_status lv_community_reset(lv_community *c);
lv_status lv_community_run(lv_community *c, const lv_params *p, lv_result *r);

/* Snapshot access for parity checks (no reference counterpart: the reference
 * reads these device_vectors directly). */
lv_status lv_community_level_info(const lv_community *c, uint32_t *n, uint64_t *E, int32_t *level);
lv_status lv_community_get_partition(const lv_community *c, uint32_t *comm, size_t n);
lv_status lv_community_set_partition(lv_community *c, const uint32_t *comm, size_t n);
lv_status lv_community_copy_graph(const lv_community *c, uint64_t *row, uint32_t *col, uint32_t *w,
                                  uint32_t *k);
/* One best-move decision from the current snapshot (no move applied):
 * target[v] = community v would move to, or comm[v].  active may be NULL (all).
 * kind LV_SWEEP_FAST requires the identity partition. */
lv_status lv_best_moves(lv_community *c, int kind, const lv_params *p, const uint8_t *active,
                        uint32_t *target, uint64_t *e_active);
/* Exact modularity parts of the current partition, from scratch on the device. */
lv_status lv_community_modularity_parts(lv_community *c, int64_t *sigma_in, uint64_t *sum_tot_sq,
                                        double *q);
/* No counterpart in the reference: lv_community_modularity_parts for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The CSR (row[n+1],
 * col[E], w[E] or NULL for unit weights), the degrees kdeg[n] and the
 * partition comm[n] are device pointers of the caller; every community id
 * must be below n (else LV_ERR_ARG); total_weight is the graph's m.  The
 * call allocates nothing: its scratch (community totals and counters) is a
 * fixed device array of the library, so n is at most 16,777,216
 * (LV_ERR_RANGE above).  stream may be NULL; the call returns after the
 * stream has drained.  Calls that overlap on one device share that array
 * and race. */
lv_status lv_modularity_parts_noalloc(uint32_t n, uint64_t E, const uint64_t *row,
                                      const uint32_t *col, const uint32_t *w,
                                      const uint32_t *kdeg, const uint32_t *comm,
                                      double total_weight, void *stream, int64_t *sigma_in,
                                      uint64_t *sum_tot_sq, double *q);
/* Tile choice of the best-move kernels: sel[2*b + f] = candidate index of
 * degree bucket b (0-7) for hash/sort (f = 0) and fast (f = 1) sweeps.  get
 * returns the tuned (or default) choice; set pins one, e.g. a table tuned
 * earlier, and marks the community tuned.  The candidates per bucket are
 * lanes per vertex {4,2,8}, {8,4,16}, {16,8,32}, {32,16,64}, {64,32},
 * {64,256}, {256}, {256}. */
lv_status lv_community_get_tiles(const lv_community *c, uint8_t sel[16]);
/* The best-move buckets' degree bounds in use (as lv_params.bucket_max_deg). */
lv_status lv_community_get_buckets(const lv_community *c, uint32_t max_deg[8]);
lv_status lv_community_set_tiles(lv_community *c, const uint8_t sel[16]);
lv_status lv_kernel_stats_get(const lv_community *c, lv_kernel_stats *s);
lv_status lv_last_run_kernel_stats(lv_kernel_stats *s);
/* Device copy bandwidth (read + write bytes per second / 1e9) of `reps`
 * launches of the plain copy kernel k_stream_copy over `bytes`: the measured
 * ceiling next to the nominal 8 TB/s, and the known-size reference for
 * FETCH_SIZE/WRITE_SIZE counters. */
lv_status lv_bandwidth_probe(uint64_t bytes, int reps, double *gbs);
/* Random-gather ceiling of a graph (no counterpart in the reference): the
 * rate, in gathers per second over `reps` launches, of a kernel that streams
 * the column ids of the community's input graph and loads arr[col[j]] for
 * each (arr: `width` = 4 or 8 bytes per vertex), the dependent access of
 * every best-move and commit pass with none of their other work. */
lv_status lv_gather_probe(const lv_community *c, uint32_t width, int reps, double *gathers_per_s);
/* Kernel name of best-move variant v for gain mode `gain` (as rocprofv3 lists it). */
lv_status lv_kernel_variant_name(int variant, int gain, char *buf, size_t len);
/* Test hook, no counterpart in the reference: on != 0 makes every commit keep
 * its next-sweep marks in the byte array, the path of graphs with 2^31 or more
 * vertices (below that, marks ride in bit 63 of each vertex's packed
 * (community, target) word); 0 restores the default.  Process-wide; set it
 * only while no run is in flight. */
lv_status lv_debug_mark_bytes(int32_t on);

/* ---- batches: many independent graphs in one launch sequence ----
 * No counterpart in the reference (it runs one graph per process, main.cu:16-22).
 * The graphs run side by side as one disjoint union on the device; each keeps
 * the reference's own level loop, thresholds and counters, and its result
 * equals lv_louvain on that graph alone.  For callers with thousands of small
 * graphs, where per-graph launches and synchronisations dominate. */
typedef struct lv_batch lv_batch;

typedef struct lv_batch_result {
    /* per graph (arrays of n_graphs, each may be NULL) */
    double *modularity;
    uint32_t *n_of_best_communities;
    int32_t *levels;
    int32_t *sweeps;
    /* best communities of every graph back to back, graph-local ids (sum of
     * the graphs' vertex counts), or NULL */
    uint32_t *best_communities;
    /* whole batch */
    int32_t union_levels;
    int32_t union_sweeps;
    double seconds;
    uint64_t traversed_entries;
} lv_batch_result;

/* The graphs are copied (host) and uploaded as one union; stream may be NULL. */
lv_status lv_batch_create(const lv_graph *const *graphs, uint32_t n_graphs, void *stream,
                          lv_batch **out);
void lv_batch_destroy(lv_batch *b);
lv_status lv_batch_info(const lv_batch *b, uint32_t *n_graphs, uint64_t *n_vertices,
                        uint64_t *n_entries);
/* Runs every graph from the identity partition (the batch can be run again).
 * Chunk emulation applies only to graphs of at most chunk_entries entries
 * (one chunk); a larger graph with chunk_emulation set is LV_ERR_ARG. */
lv_status lv_batch_louvain(lv_batch *b, const lv_params *p, lv_batch_result *r);

/* ---- served Louvain: continuous batching of submitted graphs ----
 * No counterpart in the reference.  Jobs are admitted at union-level
 * boundaries while their estimated device footprint fits hbm_budget_bytes,
 * run together on the device (one disjoint union per level, every job with
 * its own level loop), and retired when their level loop ends; a job's result
 * equals lv_louvain on its graph alone.  One host thread drives a server. */
typedef struct lv_server lv_server;

typedef struct lv_job_result {
    lv_status status;                 /* LV_OK, or why the job could not run */
    double modularity;
    uint32_t n_of_best_communities;
    uint32_t n;                       /* vertices of the job's graph */
    int32_t levels;
    int32_t sweeps;
    double seconds_queued;            /* submit -> admission */
    double seconds_total;             /* submit -> retirement */
} lv_job_result;

/* hbm_budget_bytes caps the estimated footprint of the active jobs
 * (lv_server_footprint); the upload ring (budget / 8, at most 16 GiB) and the
 * grow-only scratch of the union come on top.  A server whose step returned
 * an error is left in an unspecified state: destroy it. */
lv_status lv_server_create(const lv_params *p, uint64_t hbm_budget_bytes, void *stream,
                           lv_server **out);
void lv_server_destroy(lv_server *s);
/* The graph must stay valid until the job is retired.  A graph whose
 * footprint exceeds the budget is LV_ERR_RANGE. */
lv_status lv_server_submit(lv_server *s, const lv_graph *g, uint64_t *job_id);
/* One union level; counts after it (each may be NULL). */
lv_status lv_server_step(lv_server *s, uint32_t *n_active, uint32_t *n_queued,
                         uint32_t *n_retired);
/* Ids of retired jobs not yet taken, oldest first: up to cap are written and
 * *count says how many; the rest stay for the next call. */
lv_status lv_server_take_finished(lv_server *s, uint64_t *ids, size_t cap, size_t *count);
/* Result of a retired job; best may be NULL (else n >= the job's vertex
 * count).  LV_ERR_STATE if the job is not retired.  release != 0 frees it. */
lv_status lv_server_result(lv_server *s, uint64_t job_id, lv_job_result *r, uint32_t *best,
                           size_t n, int release);
/* Estimated device bytes of a job while it is active. */
lv_status lv_server_footprint(const lv_graph *g, uint64_t *bytes);

/* Cumulative counters of a server (host wall time per phase of its steps). */
typedef struct lv_server_stats {
    uint64_t steps;
    uint64_t admitted;
    uint64_t retired;
    uint64_t uploaded_bytes;         /* host-to-device bytes of admitted graphs */
    double seconds_assemble;         /* union assembly, uploads included */
    double seconds_sweeps;
    double seconds_aggregate;        /* renumbering, retirement, contraction */
} lv_server_stats;
lv_status lv_server_stats_get(const lv_server *s, lv_server_stats *out);

/* Seeded synthetic graph generators (host, multithreaded; counter-based PRNG,
 * so the output is independent of the thread count).  Edges are written as
 * (a[i], b[i], w[i]) with dense ids; *m receives the count.  Pass NULL arrays
 * to query the capacity needed in *m. */
/* R-MAT: 2^scale vertices, edge_factor*2^scale draws, quadrant probabilities
 * a,b,c (d = 1-a-b-c); symmetrised, self-loops and duplicates dropped, ids of
 * non-isolated vertices compacted in order. */
lv_status lv_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                      int64_t *ea, int64_t *eb, uint64_t cap, uint64_t *m, uint32_t *n_out);
/* Barabasi-Albert preferential attachment: n vertices, m_attach edges per new vertex. */
lv_status lv_gen_ba(uint32_t n, uint32_t m_attach, uint64_t seed, int64_t *ea, int64_t *eb,
                    uint64_t cap, uint64_t *m);
/* Planted partition with power-law degrees and community sizes, mixing mu,
 * integer weights uniform in [w_lo, w_hi]; membership written to planted[n]. */
lv_status lv_gen_planted(uint32_t n, double deg_exp, double mean_deg, uint32_t max_deg,
                         uint32_t n_comm, double size_exp, uint32_t size_lo, uint32_t size_hi,
                         double mu, uint32_t w_lo, uint32_t w_hi, uint64_t seed,
                         int64_t *ea, int64_t *eb, uint32_t *ew, uint64_t cap, uint64_t *m,
                         uint32_t *planted);

#ifdef __cplusplus
}
#endif
#endif
