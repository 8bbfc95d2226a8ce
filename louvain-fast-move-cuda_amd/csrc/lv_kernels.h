// lv_kernels.h -- host launchers for the gfx950 kernels in lv_kernels.hip.
//
// Data layout of one level in HBM (all arrays device resident):
//   row[n+1] u64   CSR offsets, entries of a row sorted by destination
//   col[E]   u32   destinations
//   wt[E]    u32   integer weights (absent at a unit-weight level 0)
//   kdeg[n]  u32   weighted degree k_v
//   comm[n]  u32   current community of each vertex
//   target[n] u32  community chosen by the best-move pass (== comm unless moving)
//   tot[n]   u32   community totals (sum of k over members), exact
//   bucket[n] u8   degree bucket of each vertex
//   mark[n]  u8    "re-examine next sweep" flags written by the commit pass
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lv {

constexpr int kNumBuckets = 9;        // 8 LDS-table buckets + global-table hubs
constexpr int kHubBucket = 8;
constexpr int kStatSlots = 64;        // spread of the per-block statistic atomics
constexpr int kHubParts = 32;         // community partitions per hub vertex
#ifndef LV_HUB_CHUNKS
#define LV_HUB_CHUNKS 16
#endif
// Row chunks per hub vertex, one workgroup each, in the hub kernels that walk
// the row (counting, scattering, the fast sweep): a few entries per lane, so a
// workgroup's fixed chain of loads is paid over many gathers (R-MAT-22, hash
// sweeps' hub path on one stream: 2.73 ms per run with 32 chunks, 2.61 with
// 16, 2.75 with 8; fast sweeps 0.59, 0.51, 0.53).
constexpr int kHubChunks = LV_HUB_CHUNKS;
constexpr int kBigBuckets = 2;        // large-row buckets (bucket 7, hubs) committed over movers
constexpr uint32_t kMaxChunks = 4096; // reference chunks per level (STEP_ROUND emulation)

// Maximum degree handled by each LDS bucket; larger goes to the hub bucket.
constexpr uint32_t kBucketMaxDeg[kNumBuckets - 1] = {4, 8, 16, 32, 64, 256, 1024, 4096};

// Degree bounds of the LDS buckets at run time: a row of d entries goes to the
// first bucket b with d <= max_deg[b], to the hubs above max_deg[7].  Each
// bound is at most kBucketMaxDeg[b] (so a bucket's tables and tiles hold its
// rows), except bucket 5's, which may reach kWide5MaxDeg (its table kernels
// then use tables of table_slots(kWide5MaxDeg) slots); the bounds do not
// decrease, and max_deg[7] >= 2 (a row in the hub path then has a non-self
// entry).  Results do not depend on the bounds.
constexpr uint32_t kWide5MaxDeg = 512;
struct BucketBounds {
    uint32_t max_deg[kNumBuckets - 1] = {4, 8, 16, 32, 64, 256, 1024, 4096};
};
// The sets the tuner chooses from (the first is the default).
constexpr int kBoundSets = 3;
constexpr uint32_t kBoundSet[kBoundSets][kNumBuckets - 1] = {
    {4, 8, 16, 32, 64, 256, 1024, 4096},
    {3, 6, 12, 24, 48, 192, 768, 3072},
    {2, 4, 8, 16, 32, 128, 512, 2048}};
bool bounds_valid(const uint32_t *max_deg);

// Slots of a best-move table for a row of d entries: d + d/2 + 1 rounded up
// to a multiple of 4, so the table is at most two thirds full and always has
// an empty slot (a lookup of a missing key ends).  A bucket's LDS table has
// the slots of its maximum degree; the LDS per vertex bounds how many
// vertices a CU holds at once, which bounds these latency-bound kernels.
constexpr uint32_t table_slots(uint64_t d) { return (uint32_t)((d + d / 2 + 1 + 3) & ~3ull); }
constexpr int kTableT[kNumBuckets - 1] = {
    (int)table_slots(kBucketMaxDeg[0]), (int)table_slots(kBucketMaxDeg[1]),
    (int)table_slots(kBucketMaxDeg[2]), (int)table_slots(kBucketMaxDeg[3]),
    (int)table_slots(kBucketMaxDeg[4]), (int)table_slots(kBucketMaxDeg[5]),
    (int)table_slots(kBucketMaxDeg[6]), (int)table_slots(kBucketMaxDeg[7])};
static_assert(kTableT[0] == 8 && kTableT[5] == 388 && kTableT[6] == 1540 && kTableT[7] == 6148,
              "table sizes");
// The LDS tables of bucket 7 (best moves and contraction, 8 bytes a slot)
// hold fewer slots than table_slots(4096): four workgroups then fit a CU's
// LDS instead of three.  A row or community takes min(table_slots(d),
// kTable7) slots, so those above 3,306 entries fill their table beyond two
// thirds (up to 83%; a slot stays empty).  The contraction's table stages 64
// members at a time beside it.
constexpr int kTable7 = 4960;
static_assert(kTable7 > (int)kBucketMaxDeg[7] && kTable7 * 8 + 64 * 16 + 256 <= 160 * 1024 / 4,
              "bucket-7 tables");
constexpr int kMoveTableT[kNumBuckets - 1] = {kTableT[0], kTableT[1], kTableT[2], kTableT[3],
                                              kTableT[4], kTableT[5], kTableT[6], kTable7};

// Tile (lanes per vertex) candidates of each LDS bucket's best-move kernels;
// the first is the default, 0 ends a list.  The LDS table of a bucket stays
// kTableT whatever the tile.  Few lanes per vertex (a few entries per lane)
// win in the small buckets: a wave then holds many vertices' chains.
constexpr int kTileChoices = 3;
// best-move tables of at most this many slots carry each community's total
// (PK); bucket 7's would need 72 KB of LDS per vertex packed, 48 KB unpacked
constexpr int kPackedMaxT = kTableT[6];
constexpr int kTileG[kNumBuckets - 1][kTileChoices] = {
    {2, 1, 4}, {4, 2, 8}, {8, 4, 16}, {16, 8, 32}, {32, 16, 64}, {64, 32, 256}, {256, 0, 0},
    {256, 0, 0}};
// Chosen candidate per (bucket, fast sweep or not).
struct TileChoice {
    uint8_t sel[kNumBuckets - 1][2] = {};
};

enum Stat : int {
    ST_E_ACTIVE = 0,     // non-self entries of examined vertices
    ST_CONSIDERED = 1,   // examined vertices with >= 1 non-self entry
    ST_CANDIDATES = 2,   // distinct (vertex, neighbour community) candidates
    ST_MOVED = 3,        // vertices that changed community
    ST_DSIGMA = 4,       // change of sigma_in (two's complement)
    ST_SUMSQ = 5,        // sum over communities of tot^2
    ST_TWRITES = 6,      // target writes of the best-move pass (movers)
    ST_MOVER_ENTRIES = 7,  // entries of the movers' rows, read by the commit
    kNumStats = 8
};

struct LevelView {
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *wt;       // nullptr: unit weights
    const uint32_t *kdeg;
    uint32_t n;
    uint64_t E;
};

// Hub rows (> 4096 entries) are reduced by community hash partition: hub h
// owns entries [ent_off[h], ent_off[h] + degree) of stage and pair.
struct HubTables {
    uint32_t *slot_of;        // [n] hub slot of a vertex (LV_EMPTY if not a hub)
    uint64_t *ent_off;        // [n_hubs] offset of the hub's entries in stage / pair
    unsigned long long *stage;  // [hub entries] (community << 32) | weight pairs of each
                                // chunk, packed at the start of the chunk's range
    uint32_t *ccnt;           // [n_hubs * kHubChunks] pairs each chunk wrote
    unsigned long long *pair; // [hub entries] (community << 32) | weight in partition order
    uint32_t *cnt;            // [n_hubs * kHubParts] entries per partition, zero between sweeps
    uint32_t *base;           // [n_hubs * kHubChunks * kHubParts] start of (chunk, partition)
    uint64_t *items;          // [n_hubs * kHubParts] reduce work: vertex << 32 | slot << 8 | log2 groups << 5 | group
    uint32_t *n_items;        // [1] items posted this sweep, zero between sweeps
    int64_t *part_s;          // [n_hubs * kHubParts] partial best scores (reduce groups, or
                              // the fast sweep's chunks)
    uint32_t *part_c;         // [n_hubs * kHubParts] partial best communities
    uint32_t *self;           // [n_hubs] weight to own community, zero between sweeps
};

// Several graphs swept as one disjoint union: the graph id of a vertex selects
// its graph's 2m and m, and whether that graph is still sweeping this level.
// Communities never span graphs (no entry does), so everything else is shared.
struct BatchView {
    const uint32_t *gid = nullptr;        // [n] graph of each vertex; nullptr: a single graph
    const int64_t *m2 = nullptr;          // [graphs] 2m
    const double *tw = nullptr;           // [graphs] m (the reference's total_weight)
    const uint8_t *live = nullptr;        // [graphs] 1 while the graph sweeps at this level
    unsigned long long *dsig = nullptr;   // [graphs] sigma_in change of the sweep (two's complement)
    unsigned long long *sumsq = nullptr;  // [graphs] sum of tot^2 after the sweep
};

struct SweepArgs {
    BatchView bt;
    LevelView g;
    const uint32_t *comm;
    uint32_t *target;
    uint32_t *tot;
    uint8_t *mark;
    const uint8_t *bucket;
    unsigned long long *stats;   // [kNumStats * kStatSlots] counters of commit/finalize
    unsigned long long *bstats;  // [kNumBuckets * kNumStats * kStatSlots] best-move counters per bucket
    hipEvent_t *bev;             // nullptr, or 2 events per bucket around its best-move launch
    uint32_t bev_mask = ~0u;     // buckets whose launches bev times
    int64_t m2;                  // 2m as an integer
    double tw;                   // the reference's total_weight (m)
    uint32_t v_end;              // vertices >= v_end are not examined (chunk emulation)
    int kind;                    // 0 fast, 1 sort, 2 hash
    int gain;                    // 0 exact, 1 ref_f32
    bool highest;                // ties to the highest community id
    TileChoice tiles;            // best-move kernel tile per bucket
    uint32_t *movers;            // [n] commit scratch: movers of the large-row buckets
    uint32_t *n_movers;          // [kBigBuckets] their counts
    unsigned long long *ct;      // [n] scratch: comm | tot[comm] << 32 during a hash sweep's
                                 // best-move pass, then comm | target << 32 for its commit
    bool counters_zeroed = false; // n_movers and the next-list counts were zeroed by sweep_reset
    bool wide5 = false;          // bucket 5's bound is above kBucketMaxDeg[5] (wider tables)
};

// ct <- comm | tot[comm] << 32 of every vertex: the packed operand of the hash
// sweep's table kernels.  move_pass builds it itself; the tile tuner calls it
// before timing move_bucket on a hash sweep.
void pack_comm_tot(const SweepArgs &a, hipStream_t s);

// Commits keep their marks in the byte array even below 2^31 vertices (the
// lv_debug_mark_bytes test hook).
void set_mark_bytes(bool on);

// One LDS bucket's best-move launch with tile candidate `sel` (the sweep kind
// and the bucket's list as in move_pass); used by the tile tuner.
void move_bucket(const SweepArgs &a, int bucket, int sel, const uint32_t *list, uint32_t count,
                 hipStream_t s);

// ---- primitives ----
void fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
// one launch at the start of a sweep: stats, movers counts and next-list counts to zero
void sweep_reset(unsigned long long *stats, size_t n_stats, uint32_t *n_movers,
                 uint32_t *next_counts, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s);
void fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s);
void iota_u32(uint32_t *p, size_t n, hipStream_t s);
// dst[idx[i]] = val[i]
void scatter_u32(uint32_t *dst, const uint32_t *idx, const uint32_t *val, size_t n, hipStream_t s);
// plain device copy of `bytes` (a multiple of 16) with k_stream_copy
void stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s);
// gather probe: arr[col[j]] (width 4 or 8 bytes) over E column ids on a grid
// of kGatherProbeBlocks workgroups of 256 lanes; out holds a word per lane
constexpr uint32_t kGatherProbeBlocks = 16384;
constexpr size_t kGatherProbeWords = (size_t)kGatherProbeBlocks * 256;
void gather_probe(const uint32_t *col, uint64_t E, const void *arr, int width,
                  unsigned long long *out, hipStream_t s);
// number of non-self entries of every vertex
void nonself_counts(const uint64_t *row, const uint32_t *col, uint32_t n, uint32_t *nsel,
                    hipStream_t s);
// csel[k] += non-self entries of listed vertices in chunk k (vertices [cv[k], cv[k+1]))
void chunk_selected(const uint32_t *list, uint32_t count, const uint32_t *nsel, const uint32_t *cv,
                    uint32_t nc, unsigned long long *csel, hipStream_t s);
// src[j] = the row of entry j, deg[v] = row[v+1] - row[v]
void coo_sources(const uint64_t *row, uint32_t n, uint32_t *src, uint32_t *deg, hipStream_t s);
// deg[i] = row[ids[i]+1] - row[ids[i]]
void gather_degrees(const uint64_t *row, const uint32_t *ids, uint32_t n, uint64_t *deg,
                    hipStream_t s);
// tot[c] = sum of kdeg over vertices with comm == c
void community_totals(const uint32_t *comm, const uint32_t *kdeg, uint32_t n, uint32_t *tot,
                      hipStream_t s);
size_t scan_scratch_elems(size_t n);   // elements of T scratch for scan
void scan_u32(const uint32_t *in, uint32_t *out, size_t n, bool inclusive, uint32_t *scratch,
              hipStream_t s);
void scan_u64(const uint64_t *in, uint64_t *out, size_t n, bool inclusive, uint64_t *scratch,
              hipStream_t s);
// LSD radix sort of (key, val) pairs over the low `bits` bits of key.  Result
// lands in (k0, v0); (k1, v1) are same-sized buffers; hist needs
// radix_hist_elems(n) u32.
size_t radix_hist_elems(size_t n);
void radix_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int bits,
                      uint32_t *hist, uint32_t *scan_scratch, hipStream_t s);

// ---- level setup ----
void bucket_ids(const uint64_t *row, uint32_t n, uint8_t *bucket, const BucketBounds &bounds,
                hipStream_t s);
// lists[b] receives the vertices of bucket b (all vertices); counts[b] their number.
void build_lists_all(const uint8_t *bucket, uint32_t n, uint32_t *const *lists, uint32_t *counts,
                     hipStream_t s);
// sigma_in of partition comm from scratch, and sum of tot^2 (stats slots)
void modularity_parts(const LevelView &g, const uint32_t *comm, const uint32_t *tot,
                      unsigned long long *stats, hipStream_t s);

// sigma_in and sum of k^2 of the identity partition (self-loop weights)
void identity_parts(const LevelView &g, unsigned long long *stats, hipStream_t s);

// ---- sweep ----
// Side streams of a pass whose launches are independent: the pass forks
// after `fork` is recorded on the main stream and joins before it returns.
struct SideStreams {
    hipStream_t s[3] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[3] = {};
    bool ready() const { return s[0] != nullptr; }
};

// side: when ready, buckets 5, 7 and the hubs run on side streams beside
// buckets 0-4, then bucket 6 alone (a bucket's profiling events on its own
// stream).
void move_pass(const SweepArgs &a, const HubTables &hub, uint32_t *const *lists,
               const uint32_t *counts, hipStream_t s, const SideStreams *side = nullptr);
// side: when ready, buckets 0-5 commit on a side stream beside the large rows.
void commit_pass(const SweepArgs &a, uint32_t *const *lists, const uint32_t *counts,
                 hipStream_t s, const SideStreams *side = nullptr);
// comm := target for movers, sum tot^2, compact marked vertices into next lists
void finalize_pass(const SweepArgs &a, uint32_t *const *next_lists, uint32_t *next_counts,
                   hipStream_t s);

// ---- batches of graphs ----
// counts[gid[c]] += 1 for every non-empty community c
void batch_comm_counts(const uint32_t *tot, const uint32_t *gid, uint32_t n,
                       unsigned long long *counts, hipStream_t s);
// gid_new[map[c]-1] = gid[c] for every non-empty community c
void batch_next_gid(const uint32_t *tot, const uint32_t *map, const uint32_t *gid, uint32_t n,
                    uint32_t *gid_new, hipStream_t s);

// ---- aggregation ----
// new_id[c] = map[c]-1 for communities with tot>0; n_new via *map_last (device)
void renumber_flags(const uint32_t *tot, uint32_t n, uint32_t *flags, hipStream_t s);
void update_best(uint32_t *best, uint32_t n0, const uint32_t *comm, const uint32_t *map,
                 hipStream_t s);
void new_degrees(const uint32_t *tot, const uint32_t *map, uint32_t n, uint32_t *k_new,
                 hipStream_t s);

// ---- aggregation by community hashing (lv_aggregate.hip) ----
struct AggArgs {
    LevelView g;
    const uint32_t *newc;      // [n]
    const uint32_t *members;   // [n] vertices sorted by new community
    const uint32_t *mstart;    // [n_new + 1]
    const uint64_t *P;         // [n + 1] prefix of member degrees
    const uint64_t *crow;      // [n_new + 1] virtual entry offsets of communities
    uint32_t *tcol;            // [E] staging
    uint32_t *tw;              // [E] staging
    uint32_t *rcount;          // [n_new], zeroed
    // The next level's first sweep -- a fast sweep at the identity partition
    // over the contracted rows -- decided while the rows are built (dgain: 0
    // exact gains, 1 the reference's f32 rounding; -1 not asked).  Every new
    // vertex c gets dtarget[c] (its best neighbour under the fast sweep's
    // rule, or c), and the sweep's best-move statistics go to dstats.
    int dgain = -1;
    int64_t m2 = 0;                                 // 2m of the graph
    double tw_total = 0;                            // m
    const unsigned long long *newck = nullptr;      // [n] newc[v] | new degree of newc[v] << 32
    const uint32_t *knew = nullptr;                 // [n_new] new degrees
    uint32_t *dtarget = nullptr;                    // [n_new]
    unsigned long long *dstats = nullptr;           // one statistics section, zeroed
    BatchView dbt;   // batches: the new vertices' graphs (their 2m, m and live flags)
};
// newc[v] = map[comm[v]]-1; key/val prepared for the vertex sort (key = newc,
// val = v); with newck, also newc[v] | tot[comm[v]] << 32 (tot[comm[v]] is the
// new degree of newc[v])
void agg_newc(const uint32_t *comm, const uint32_t *map, uint32_t n, uint32_t *newc, uint64_t *key,
              uint32_t *val, const uint32_t *tot, unsigned long long *newck, hipStream_t s);
// member starts (mstart[n_new] = n) and member degrees (mdeg[n] = 0)
void agg_members(const uint64_t *skey, const uint32_t *members, const uint64_t *row, uint32_t n,
                 uint32_t n_new, uint32_t *mstart, uint64_t *mdeg, hipStream_t s);
void agg_comm_rows(const uint64_t *P, const uint32_t *mstart, uint32_t n_new, uint64_t *crow,
                   hipStream_t s);
// LDS-table communities (buckets 0-7): sorted distinct rows into the staging area
void agg_tables(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts, hipStream_t s);
// Giant communities (above 4096 virtual entries): 2^lgp destination hash
// partitions (one per kGiantPartEntries entries, at most 2^kGiantMaxPartBits)
// for the LDS reduce, and tiles of 2^lgt entries (at least 16 per partition)
// for the gather and scatter.
constexpr uint64_t kGiantPartEntries = 4096;
constexpr int kGiantMaxPartBits = 12;
constexpr int kGiantMinTileBits = 12;
struct GiantInfo {
    uint64_t off;     // first entry of the community in the pair buffers
    uint64_t mbase;   // first of its (tile, partition) run starts
    uint32_t c;       // community id
    uint32_t pbase;   // first of its partition counters
    uint32_t lgp;     // log2 of its partition count
    uint32_t lgt;     // log2 of its tile size
};
struct GiantArgs {
    const GiantInfo *info;          // [giants]
    const uint64_t *tiles;          // [n_tiles] giant index << 32 | tile
    uint32_t n_tiles;
    const uint64_t *items;          // [n_items] giant index << 32 | partition
    uint32_t n_items;
    unsigned long long *p0;         // [entries] pairs of each tile at its range's start, then
                                    // distinct pairs
    uint32_t *tcnt;                 // [n_tiles] pairs each tile wrote
    unsigned long long *p1;         // [entries] pairs in partition order
    uint32_t *pcnt;                 // [partitions] entries per partition, zeroed
    uint32_t *poff;                 // [partitions] exclusive scan of pcnt: a partition's pairs
    uint32_t *dcnt;                 // [partitions] distinct destinations per partition
    uint32_t *doff;                 // [partitions] exclusive scan of dcnt
    uint32_t *base;                 // [(tile, partition) runs] start of a tile's run
    uint32_t *scan_tmp;             // scan scratch for the partition counts
    uint32_t max_parts;             // largest partition count of the plan (LDS per workgroup)
    uint32_t n_giants;
    // with AggArgs::dgain >= 0: each partition's best candidate and non-self
    // destinations, reduced per giant afterwards
    int64_t *dpart_s;               // [partitions]
    uint32_t *dpart_c;              // [partitions]
    uint32_t *dpart_n;              // [partitions]
};
// distinct (destination, weight) pairs of every giant community and its
// rcount; then, once the new offsets are known, its row copied to place
void agg_giant(const AggArgs &x, const GiantArgs &g, uint32_t n_parts, hipStream_t s);
void agg_giant_copy(const AggArgs &x, const GiantArgs &g, uint32_t n_parts, const uint64_t *out_row,
                    uint32_t *col, uint32_t *w, hipStream_t s);
void heads_u64(const uint64_t *keys, uint64_t n, uint32_t *heads, hipStream_t s);
// runs of sorted keys: idx = inclusive scan of heads, S = inclusive scan of weights
void agg_runs(const uint64_t *keys, const uint32_t *vals, const uint32_t *idx, const uint64_t *S,
              uint64_t n, uint64_t *ukey, uint64_t *ws, uint64_t *we, hipStream_t s);
// rows of the LDS-table communities copied to their final offsets
void agg_copy(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts,
              const uint64_t *out_row, uint32_t *col, uint32_t *w, hipStream_t s);
void set_u64(uint64_t *p, uint64_t v, hipStream_t s);
void widen_u32(const uint32_t *in, uint32_t n, uint64_t *out, hipStream_t s);
// per-bucket sizes of a degree array given by offsets row[n+1]
void bucket_counts(const uint8_t *bucket, uint32_t n, uint32_t *counts, hipStream_t s);

}  // namespace lv
