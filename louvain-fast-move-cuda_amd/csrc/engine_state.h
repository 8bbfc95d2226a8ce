// engine_state.h -- per-level working state of the sweep loop (degree
// buckets, active lists, hub tables, statistics).  Internal to the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/device_array.h"
#include "../../include/graph.h"
#include "lv_kernels.h"

struct Community;

namespace lv {

// best-move kernel variants timed one by one: (bucket 0-7, fast, weighted, tile)
constexpr int kMoveVariants = 96;
inline int move_variant(int bucket, bool fast, bool weighted, int tile) {
    return ((bucket * 2 + (fast ? 1 : 0)) * 2 + (weighted ? 1 : 0)) * kTileChoices + tile;
}

struct KernelStatsAcc {
    uint64_t move_launches = 0;
    double move_ms = 0;
    uint64_t move_bytes = 0;
    uint64_t move_entries = 0;
    double sweep_ms = 0;
    uint64_t sweeps = 0;
    uint64_t var_launches[kMoveVariants] = {};
    double var_ms[kMoveVariants] = {};
    uint64_t var_bytes[kMoveVariants] = {};
    uint64_t var_entries[kMoveVariants] = {};
    // per degree bucket (8 = hubs), [fast sweep?]: launches, event time,
    // algorithmic bytes, entries
    uint64_t bk_launches[kNumBuckets][2] = {};
    double bk_ms[kNumBuckets][2] = {};
    uint64_t bk_bytes[kNumBuckets][2] = {};
    uint64_t bk_entries[kNumBuckets][2] = {};
    // commit passes and contractions
    uint64_t commit_passes = 0;
    double commit_ms = 0;
    uint64_t commit_bytes = 0, commit_entries = 0;
    uint64_t contractions = 0;
    double contract_ms = 0;
    uint64_t contract_bytes = 0, contract_entries = 0;
};

extern KernelStatsAcc g_last_run_stats;

// stats layout: one section of kNumStats*kStatSlots counters per degree bucket
// (best-move kernels) and a last, global section (commit, finalize,
// modularity); host_stats mirrors it and then holds kNumBuckets u32 list sizes.
constexpr size_t kSection = (size_t)kNumStats * kStatSlots;
constexpr size_t kGlobalSection = (size_t)kNumBuckets * kSection;
constexpr size_t kStatWordsOffset = (size_t)(kNumBuckets + 1) * kSection;
constexpr size_t kHostWords = kStatWordsOffset + kNumBuckets;

// Grow-only device scratch: one buffer per slot, reused across levels and
// runs so that the level loop does not call hipMalloc/hipFree (which
// synchronise the device) once the largest level has been seen.
struct Workspace {
    std::vector<DeviceArray<uint8_t>> slots;
    template <typename T>
    T *get(size_t slot, size_t n) {
        if (slots.size() <= slot) slots.resize(slot + 1);
        const size_t bytes = (n ? n : 1) * sizeof(T);
        if (slots[slot].size() < bytes) {
            slots[slot].release();
            slots[slot].resize(bytes);
        }
        return reinterpret_cast<T *>(slots[slot].data());
    }
};

constexpr size_t kSetupSlot = 40;   // workspace slots used by level setup
constexpr size_t kSortSlot = 60;    // workspace slots used by the sort sweep
constexpr size_t kChunkSlot = 90;   // workspace slots used by chunk emulation

struct ChunkKey {
    const void *row = nullptr;
    uint32_t n = 0;
    uint64_t E = 0, step = 0;
    bool operator==(const ChunkKey &o) const {
        return row == o.row && n == o.n && E == o.E && step == o.step;
    }
};

struct SweepState {
    uint32_t n = 0;
    DeviceArray<uint8_t> bucket;
    DeviceArray<uint8_t> mark;
    DeviceArray<uint32_t> target;
    DeviceArray<uint32_t> movers;               // commit scratch of the large-row buckets
    DeviceArray<uint32_t> n_movers;
    DeviceArray<unsigned long long> ct;         // commit scratch: comm | target << 32
    DeviceArray<uint32_t> list_buf[2];
    DeviceArray<uint32_t> counts_dev;
    // statistics (kStatWordsOffset words) followed by the next lists' sizes
    // (kNumBuckets u32), the same layout as host_stats: one copy per sweep
    DeviceArray<unsigned long long> stats;
    uint32_t *next_counts() { return reinterpret_cast<uint32_t *>(stats.data() + kStatWordsOffset); }
    unsigned long long *host_stats = nullptr;   // pinned
    uint64_t bucket_off[kNumBuckets] = {};
    uint32_t bucket_size[kNumBuckets] = {};
    uint32_t count[kNumBuckets] = {};           // sizes of the current lists
    uint32_t *list_ptr[2][kNumBuckets] = {};
    int cur = 0;
    // the current lists hold every vertex of the level by bucket (built by
    // build_static_lists since the level's buckets and no sweep has applied
    // moves since): a level's first sweep then needs no rebuild
    bool lists_static = false;
    // hubs
    uint32_t n_hubs = 0;
    DeviceArray<uint32_t> hub_slot_of;
    DeviceArray<uint64_t> hub_ent_off;
    DeviceArray<unsigned long long> hub_stage;
    DeviceArray<uint32_t> hub_ccnt;
    DeviceArray<unsigned long long> hub_pair;
    DeviceArray<uint32_t> hub_cnt;
    DeviceArray<uint32_t> hub_base;
    DeviceArray<uint64_t> hub_items;
    DeviceArray<uint32_t> hub_n_items;
    DeviceArray<int64_t> hub_part_s;
    DeviceArray<uint32_t> hub_part_c;
    DeviceArray<uint32_t> hub_self;
    std::vector<uint64_t> hub_eoff_host;        // source of the async hub_ent_off upload
    // host plan of the last contraction's giant communities (sources of async uploads)
    std::vector<GiantInfo> giant_info;
    std::vector<uint64_t> giant_tiles, giant_items;
    Workspace ws;
    // chunk emulation: level the bounds were computed for, chunk start vertices
    ChunkKey chunk_key;
    std::vector<uint32_t> chunk_cv;
    // two reusable storages for contracted levels (level l+1 is built while
    // level l, held in the other one or in the resident level-0 graph, is read)
    GraphDevice level_buf[2];
    int next_level_buf = 0;
    // the contraction's row copy runs on side stream 0; its join event is
    // waited for by the main stream at the end of the contraction, or, when
    // the caller defers it (AggregationPhase::run), after the next level's
    // setup, which reads the new row offsets only (join_row_copy)
    bool defer_copy_join = false;
    bool row_copy_pending = false;
    // The first sweep of the next level decided by the contraction (see
    // AggArgs::dgain): asked with first_sweep_gain >= 0 before a contraction;
    // the decisions and that sweep's best-move statistics are kept for the
    // sweep loop of level pre_level, which uses them if its first sweep is a
    // fast sweep over every vertex with the same gain mode.
    int first_sweep_gain = -1;
    BatchView first_sweep_bt;   // batches: the new level's graph ids, 2m, m, live flags
    bool pre_valid = false;
    int32_t pre_level = -1;
    int pre_gain = -1;
    DeviceArray<uint32_t> pre_target;
    DeviceArray<unsigned long long> pre_stats;
    // tiles of the best-move kernels (tuned once per community when asked)
    TileChoice tiles;
    bool tuned[2] = {false, false};   // [fast sweep]
    // degree bounds of the best-move buckets (chosen by the caller or tuned
    // once per community)
    BucketBounds bounds;
    BucketBounds tuned_bounds;
    bool bounds_tuned = false;
    // side streams of the sweep passes and the contraction
    SideStreams side;
    const SideStreams *side_streams();
    // profiling
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // [3], [4]: contraction
    hipEvent_t bev[2 * kNumBuckets] = {};
    KernelStatsAcc acc;

    SweepState() {
        for (auto &e : ev) LV_HIP(hipEventCreate(&e));
        for (auto &e : bev) LV_HIP(hipEventCreate(&e));
        (void)side_streams();   // created with the community, not inside a timed run
    }
    ~SweepState();
    SweepState(const SweepState &) = delete;
    SweepState &operator=(const SweepState &) = delete;

    void setup_level(const GraphDevice &g, hipStream_t s);
    void build_static_lists(hipStream_t s);
    uint32_t *const *lists(int which);
    HubTables hub_tables();
};

// Per-graph state of a batch run as one disjoint union (csrc/batch.cpp).
struct BatchState {
    uint32_t n_graphs = 0;
    std::vector<uint32_t> voff;        // [G+1] level-0 vertex offsets of the graphs
    std::vector<uint64_t> entries;     // [G] level-0 entries of each graph
    std::vector<double> tw;            // [G] m
    std::vector<int64_t> m2;           // [G] 2m as an integer
    std::vector<int64_t> sigma0;       // [G] identity-partition sigma_in (self-loop weight)
    std::vector<uint64_t> sumsq0;      // [G] identity-partition sum of k^2
    DeviceArray<uint32_t> gid0;        // [n0] graph of each level-0 vertex (resident)
    DeviceArray<uint32_t> gid_buf[2];  // graph ids of contracted levels
    const uint32_t *gid = nullptr;     // graph ids of the current level
    int next_gid = 0;
    DeviceArray<int64_t> m2_dev;
    DeviceArray<double> tw_dev;
    DeviceArray<uint8_t> live_dev;
    DeviceArray<unsigned long long> acc;       // [3G]: dsig, sumsq, community counts
    unsigned long long *host_acc = nullptr;    // pinned [3G]
    uint8_t *host_live = nullptr;              // pinned [G]

    BatchState() = default;
    ~BatchState();
    BatchState(const BatchState &) = delete;
    BatchState &operator=(const BatchState &) = delete;

    BatchView view() {
        BatchView v;
        v.gid = gid;
        v.m2 = m2_dev.data();
        v.tw = tw_dev.data();
        v.live = live_dev.data();
        v.dsig = acc.data();
        v.sumsq = acc.data() + n_graphs;
        return v;
    }
};

// renumbering of the non-empty communities (device map, 1-based), with
// best_communities composed with it; n_new receives the community count
uint32_t *renumber(::Community &c, uint32_t &n_new);

LevelView level_view(const GraphDevice &g);
// The sort variant of the best-move pass (lv_sort_sweep.hip).
void sort_sweep(const SweepArgs &x, uint32_t *const *lists, const uint32_t *counts,
                const uint64_t *bucket_off, Workspace &ws, hipStream_t s);
uint64_t sum_slots(const unsigned long long *h, int stat);            // global section
// Sets the best-move buckets' degree bounds of a run from lv_params (or the
// tuner, on the community's first run with autotune) and rebuilds the level's
// buckets if they change.
void apply_bucket_bounds(Community &C, const lv_params &p);
uint64_t bucket_slots(const unsigned long long *h, int bucket, int stat);
uint64_t all_bucket_slots(const unsigned long long *h, int stat);
// pre_decided: the best-move pass is replaced by the decisions and statistics
// the contraction kept (SweepState::pre_target, pre_stats)
void launch_sweep(::Community &c, const lv_params &p, int kind, bool apply, uint32_t v_end,
                  bool pre_decided = false);
// Times each LDS bucket's tile candidates on the current lists (once per
// community and sweep class) and keeps the fastest in state->tiles.
void autotune_tiles(::Community &c, const lv_params &p, int kind, uint32_t v_end);

}  // namespace lv
