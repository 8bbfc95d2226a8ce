// engine.cpp -- host driver of the level loop: GraphDevice upload, Community
// state, OptimizationPhase (sweep loop), AggregationPhase (renumber +
// contract) and ModularityAlgorithms::Louvain (level loop).
//
// Reference: graph.cuh:46-99, community.h:31-106, phaseOptimization.cuh:50-206,
// phaseOptimizationUpdate.cu:135-222, phaseAggregation.cuh:19-190,
// modularityAlgorithms.h:14-81.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/community.h"
#include "../../include/modularityAlgorithms.h"
#include "../../include/phaseAggregation.h"
#include "../../include/phaseOptimization.h"
#include "../../include/phaseRefine.h"
#include "engine_state.h"
#include "lv_kernels.h"

namespace lv {

static uint64_t fnv1a_u32(const uint32_t *a, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t x = a[i];
        for (int b = 0; b < 4; ++b) {
            h ^= (x >> (8 * b)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    return h;
}

KernelStatsAcc g_last_run_stats;

void SweepState::setup_level(const GraphDevice &g, hipStream_t s) {
    n = g.n_nodes;
    lists_static = false;
    bucket.ensure(std::max<size_t>(n, 1));
    mark.ensure(std::max<size_t>(n, 1));
    target.ensure(std::max<size_t>(n, 1));
    movers.ensure(std::max<size_t>(n, 1));
    ct.ensure(std::max<size_t>(n, 1));
    n_movers.ensure(kBigBuckets);
    list_buf[0].ensure(std::max<size_t>(n, 1));
    list_buf[1].ensure(std::max<size_t>(n, 1));
    counts_dev.ensure(2 * kNumBuckets);
    stats.ensure(kHostWords);
    if (!host_stats) {
        LV_HIP(hipHostMalloc((void **)&host_stats, kHostWords * sizeof(unsigned long long),
                             hipHostMallocDefault));
    }
    fill_u8(mark.data(), 0, n, s);
    // bucket ids and per-bucket sizes (static: degrees do not change within a level)
    bucket_ids(g.neighboorhood_sum.data(), n, bucket.data(), bounds, s);
    bucket_counts(bucket.data(), n, counts_dev.data(), s);
    uint32_t cnt[kNumBuckets];
    LV_HIP(hipMemcpyAsync(host_stats, counts_dev.data(), kNumBuckets * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    wait_stream(s);
    std::memcpy(cnt, host_stats, sizeof(cnt));
    uint64_t off = 0;
    for (int b = 0; b < kNumBuckets; ++b) {
        bucket_size[b] = cnt[b];
        bucket_off[b] = off;
        off += cnt[b];
    }
    // hub tables
    n_hubs = bucket_size[kHubBucket];
    // hub work items carry the slot in 24 bits (2m < 2^31 keeps hubs < 2^19)
    if (n_hubs >= (1u << 24)) throw Error(LV_ERR_RANGE, "too many hub vertices");
    if (n_hubs > 0) {
        build_static_lists(s);
        std::vector<uint64_t> hdeg(n_hubs);
        uint64_t *hdeg_dev = ws.get<uint64_t>(kSetupSlot, n_hubs);
        gather_degrees(g.neighboorhood_sum.data(), list_buf[0].data() + bucket_off[kHubBucket],
                       n_hubs, hdeg_dev, s);
        LV_HIP(hipMemcpyAsync(hdeg.data(), hdeg_dev, n_hubs * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, s));
        wait_stream(s);
        std::vector<uint64_t> &eoff = hub_eoff_host;   // kept: read by the async upload below
        eoff.resize(n_hubs);
        uint64_t hub_entries = 0;
        for (uint32_t i = 0; i < n_hubs; ++i) {
            eoff[i] = hub_entries;
            hub_entries += hdeg[i];
        }
        hub_slot_of.ensure(n);
        fill_u32(hub_slot_of.data(), LV_EMPTY, n, s);
        hub_ent_off.ensure(n_hubs);
        hub_stage.ensure(hub_entries);
        hub_ccnt.ensure((size_t)n_hubs * kHubChunks);
        hub_pair.ensure(hub_entries);
        hub_cnt.ensure((size_t)n_hubs * kHubParts);
        hub_base.ensure((size_t)n_hubs * kHubChunks * kHubParts);
        hub_items.ensure((size_t)n_hubs * kHubParts);
        hub_n_items.ensure(1);
        fill_u32(hub_n_items.data(), 0u, 1, s);
        hub_part_s.ensure((size_t)n_hubs * kHubParts);
        hub_part_c.ensure((size_t)n_hubs * kHubParts);
        hub_self.ensure(n_hubs);
        fill_u32(hub_cnt.data(), 0u, (size_t)n_hubs * kHubParts, s);
        fill_u32(hub_self.data(), 0u, n_hubs, s);
        LV_HIP(hipMemcpyAsync(hub_ent_off.data(), eoff.data(), n_hubs * sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
        // slot_of[hubs[i]] = i, from the hub list already on the device
        uint32_t *hub_slots = ws.get<uint32_t>(kSetupSlot + 2, n_hubs);
        iota_u32(hub_slots, n_hubs, s);
        scatter_u32(hub_slot_of.data(), list_buf[0].data() + bucket_off[kHubBucket], hub_slots,
                    n_hubs, s);
    }
}

void SweepState::build_static_lists(hipStream_t s) {
    uint32_t *lp[kNumBuckets];
    for (int b = 0; b < kNumBuckets; ++b) lp[b] = list_buf[0].data() + bucket_off[b];
    build_lists_all(bucket.data(), n, lp, counts_dev.data(), s);
    cur = 0;
    for (int b = 0; b < kNumBuckets; ++b) count[b] = bucket_size[b];
    lists_static = true;
}

uint32_t *const *SweepState::lists(int which) {
    for (int b = 0; b < kNumBuckets; ++b) list_ptr[which][b] = list_buf[which].data() + bucket_off[b];
    return list_ptr[which];
}

HubTables SweepState::hub_tables() {
    HubTables h;
    h.slot_of = hub_slot_of.data();
    h.ent_off = hub_ent_off.data();
    h.stage = hub_stage.data();
    h.ccnt = hub_ccnt.data();
    h.pair = hub_pair.data();
    h.cnt = hub_cnt.data();
    h.base = hub_base.data();
    h.items = hub_items.data();
    h.n_items = hub_n_items.data();
    h.part_s = hub_part_s.data();
    h.part_c = hub_part_c.data();
    h.self = hub_self.data();
    return h;
}

const SideStreams *SweepState::side_streams() {
    if (!side.ready()) {
        for (int i = 0; i < 3; ++i) {
            LV_HIP(hipStreamCreateWithFlags(&side.s[i], hipStreamNonBlocking));
            LV_HIP(hipEventCreateWithFlags(&side.join[i], hipEventDisableTiming));
        }
        LV_HIP(hipEventCreateWithFlags(&side.fork, hipEventDisableTiming));
    }
    return &side;
}

SweepState::~SweepState() {
    for (int i = 0; i < 3; ++i) {
        if (side.join[i]) (void)hipEventDestroy(side.join[i]);
        if (side.s[i]) (void)hipStreamDestroy(side.s[i]);
    }
    if (side.fork) (void)hipEventDestroy(side.fork);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    for (auto &e : bev) if (e) (void)hipEventDestroy(e);
    if (host_stats) (void)hipHostFree(host_stats);
}

LevelView level_view(const GraphDevice &g) {
    LevelView v;
    v.row = g.neighboorhood_sum.data();
    v.col = g.edge_destination.data();
    v.wt = g.unit_weights ? nullptr : g.weights.data();
    v.kdeg = g.tot_weight_per_nodes.data();
    v.n = g.n_nodes;
    v.E = g.n_entries;
    return v;
}

uint64_t sum_slots(const unsigned long long *h, int stat) {
    uint64_t s = 0;
    for (int i = 0; i < kStatSlots; ++i) s += h[kGlobalSection + stat * kStatSlots + i];
    return s;
}

uint64_t bucket_slots(const unsigned long long *h, int bucket, int stat) {
    uint64_t s = 0;
    for (int i = 0; i < kStatSlots; ++i) s += h[(size_t)bucket * kSection + stat * kStatSlots + i];
    return s;
}

uint64_t all_bucket_slots(const unsigned long long *h, int stat) {
    uint64_t s = 0;
    for (int b = 0; b < kNumBuckets; ++b) s += bucket_slots(h, b, stat);
    return s;
}

}  // namespace lv

using namespace lv;

// ------------------------------------------------------------- GraphDevice

GraphDevice::GraphDevice(const GraphHost &g, hipStream_t s) {
    n_nodes = g.n_nodes;
    n_links = g.n_links;
    total_weight = g.total_weight;
    n_entries = g.col.size();
    unit_weights = g.unit_weights;
    neighboorhood_sum.resize(g.row.size());
    edge_destination.resize(std::max<size_t>(n_entries, 1));
    tot_weight_per_nodes.resize(std::max<size_t>(n_nodes, 1));
    std::vector<uint32_t> k(n_nodes);
    uint64_t m2w = 0;
    for (uint32_t v = 0; v < n_nodes; ++v) {
        uint64_t s2 = 0;
        if (unit_weights) s2 = g.row[v + 1] - g.row[v];
        else
            for (uint64_t j = g.row[v]; j < g.row[v + 1]; ++j) s2 += g.weights[j];
        if (s2 >= 0x80000000ull) throw Error(LV_ERR_RANGE, "vertex degree >= 2^31");
        k[v] = (uint32_t)s2;
        m2w += s2;
    }
    m2 = m2w;
    if (m2 >= 0x80000000ull && !g.disjoint_union)
        throw Error(LV_ERR_RANGE, "total weight 2m >= 2^31");
    LV_HIP(hipMemcpyAsync(neighboorhood_sum.data(), g.row.data(), g.row.size() * sizeof(uint64_t),
                          hipMemcpyHostToDevice, s));
    if (n_entries)
        LV_HIP(hipMemcpyAsync(edge_destination.data(), g.col.data(), n_entries * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    if (!unit_weights) {
        weights.resize(n_entries);
        LV_HIP(hipMemcpyAsync(weights.data(), g.weights.data(), n_entries * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    }
    if (n_nodes)
        LV_HIP(hipMemcpyAsync(tot_weight_per_nodes.data(), k.data(), n_nodes * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
    wait_stream(s);
}

// --------------------------------------------------------------- Community

Community::Community(GraphHost &g, hipStream_t stream) { init(g, stream); }

Community::Community(hipStream_t stream) {
    if (stream) {
        streams[0] = stream;
    } else {
        LV_HIP(hipStreamCreateWithFlags(&streams[0], hipStreamNonBlocking));
        own_stream_ = true;
    }
    streams[1] = streams[0];
    state = std::make_unique<SweepState>();
}
Community::Community(const GraphHost &g, hipStream_t stream) { init(g, stream); }

Community::Community(GraphDevice &&g0, hipStream_t stream) {
    if (stream) {
        streams[0] = stream;
    } else {
        LV_HIP(hipStreamCreateWithFlags(&streams[0], hipStreamNonBlocking));
        own_stream_ = true;
    }
    streams[1] = streams[0];
    graph0 = std::move(g0);
    init_resident(streams[0]);
}

void Community::init(const GraphHost &g, hipStream_t stream) {
    if (stream) {
        streams[0] = stream;
    } else {
        LV_HIP(hipStreamCreateWithFlags(&streams[0], hipStreamNonBlocking));
        own_stream_ = true;
    }
    streams[1] = streams[0];
    graph0 = GraphDevice(g, streams[0]);
    init_resident(streams[0]);
}

void Community::init_resident(hipStream_t stream) {
    (void)stream;
    graph = GraphDevice::view_of(graph0);
    n_original = graph.n_nodes;
    n_of_best_communities = graph.n_nodes;
    best_communities.resize(std::max<size_t>(graph.n_nodes, 1));
    iota_u32(best_communities.data(), graph.n_nodes, streams[0]);   // community.h:40-41
    state = std::make_unique<SweepState>();
    level = 0;
    update();
}

GraphDevice GraphDevice::view_of(const GraphDevice &g) {
    GraphDevice v;
    v.n_nodes = g.n_nodes;
    v.n_links = g.n_links;
    v.total_weight = g.total_weight;
    v.n_entries = g.n_entries;
    v.m2 = g.m2;
    v.unit_weights = g.unit_weights;
    v.neighboorhood_sum = DeviceArray<uint64_t>::borrow(g.neighboorhood_sum);
    v.edge_destination = DeviceArray<uint32_t>::borrow(g.edge_destination);
    v.weights = DeviceArray<uint32_t>::borrow(g.weights);
    v.tot_weight_per_nodes = DeviceArray<uint32_t>::borrow(g.tot_weight_per_nodes);
    return v;
}

void GraphDevice::build_coo(hipStream_t s) {
    edge_source.resize(std::max<uint64_t>(n_entries, 1));
    n_of_neighboor.resize(std::max<uint32_t>(n_nodes, 1));
    coo_sources(neighboorhood_sum.data(), n_nodes, edge_source.data(), n_of_neighboor.data(), s);
    LV_HIP(hipGetLastError());
}

void Community::reset() {
    graph = GraphDevice::view_of(graph0);
    state->next_level_buf = 0;
    state->pre_valid = false;   // decisions kept for a level of an earlier run
    level = 0;
    n_of_best_communities = graph.n_nodes;
    iota_u32(best_communities.data(), graph.n_nodes, stream());
    update();
}

Community::~Community() {
    batch.reset();
    state.reset();
    if (own_stream_ && streams[0]) (void)hipStreamDestroy(streams[0]);
}

Community::Community(Community &&o) noexcept
    : graph(std::move(o.graph)),
      communities(std::move(o.communities)),
      communities_weight(std::move(o.communities_weight)),
      best_communities(std::move(o.best_communities)),
      n_of_best_communities(o.n_of_best_communities),
      modularity(o.modularity),
      sigma_in(o.sigma_in),
      sum_tot_sq(o.sum_tot_sq),
      level(o.level),
      n_original(o.n_original),
      state(std::move(o.state)),
      batch(std::move(o.batch)),
      graph0(std::move(o.graph0)),
      own_stream_(o.own_stream_) {
    streams[0] = o.streams[0];
    streams[1] = o.streams[1];
    o.streams[0] = o.streams[1] = nullptr;
    o.own_stream_ = false;
}

Community &Community::operator=(Community &&o) noexcept {
    if (this != &o) {
        batch.reset();
        state.reset();
        if (own_stream_ && streams[0]) (void)hipStreamDestroy(streams[0]);
        graph = std::move(o.graph);
        communities = std::move(o.communities);
        communities_weight = std::move(o.communities_weight);
        best_communities = std::move(o.best_communities);
        n_of_best_communities = o.n_of_best_communities;
        modularity = o.modularity;
        sigma_in = o.sigma_in;
        sum_tot_sq = o.sum_tot_sq;
        level = o.level;
        n_original = o.n_original;
        state = std::move(o.state);
        batch = std::move(o.batch);
        graph0 = std::move(o.graph0);
        own_stream_ = o.own_stream_;
        streams[0] = o.streams[0];
        streams[1] = o.streams[1];
        o.streams[0] = o.streams[1] = nullptr;
        o.own_stream_ = false;
    }
    return *this;
}

void Community::update_best(const unsigned int *map) {
    lv::update_best(best_communities.data(), n_original, communities.data(), map, stream());
}

void Community::update() { update_level(false, 0, 0); }

void Community::update(int64_t sigma_in_carry, uint64_t sum_tot_sq_carry) {
    update_level(true, sigma_in_carry, sum_tot_sq_carry);
}

void Community::update_level(bool carry, int64_t sigma_in_carry, uint64_t sum_tot_sq_carry) {
    const uint32_t n = graph.n_nodes;
    hipStream_t s = stream();
    communities.ensure(std::max<size_t>(n, 1));
    communities_weight.ensure(std::max<size_t>(n, 1));
    iota_u32(communities.data(), n, s);                               // community.h:58-60
    if (n)
        LV_HIP(hipMemcpyAsync(communities_weight.data(), graph.tot_weight_per_nodes.data(),
                              n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    state->setup_level(graph, s);
    iota_u32(state->target.data(), n, s);
    if (carry) {
        sigma_in = sigma_in_carry;
        sum_tot_sq = sum_tot_sq_carry;
        modularity = q_from_parts();
        return;
    }
    // identity partition of a level-0 graph (rows sorted by destination):
    // sigma_in = self-loop weight, sum of tot^2 = sum of k^2
    fill_u64(state->stats.data(), 0ull, state->stats.size(), s);
    identity_parts(level_view(graph), state->stats.data() + kGlobalSection, s);
    LV_HIP(hipMemcpyAsync(state->host_stats, state->stats.data(),
                          kStatWordsOffset * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    sigma_in = (int64_t)sum_slots(state->host_stats, ST_DSIGMA);
    sum_tot_sq = sum_slots(state->host_stats, ST_SUMSQ);
    modularity = q_from_parts();
}

double Community::q_from_parts() const {
    const double tw = graph.total_weight;
    return (double)sigma_in / (2 * tw) - (double)sum_tot_sq / (4 * tw * tw);  // community.h:104
}

void Community::compute_modularity() {
    hipStream_t s = stream();
    fill_u64(state->stats.data(), 0ull, state->stats.size(), s);
    modularity_parts(level_view(graph), communities.data(), communities_weight.data(),
                     state->stats.data() + kGlobalSection, s);
    LV_HIP(hipMemcpyAsync(state->host_stats, state->stats.data(),
                          kStatWordsOffset * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    sigma_in = (int64_t)sum_slots(state->host_stats, ST_DSIGMA);
    sum_tot_sq = sum_slots(state->host_stats, ST_SUMSQ);
    modularity = q_from_parts();
}

// ------------------------------------------------------- OptimizationPhase

int OptimizationPhase::sweep_kind(int mode, int64_t exec) {
    // phaseOptimization.cuh:129-145
    if (exec == 0) return LV_SWEEP_FAST;
    if (mode == HASH || mode == ADAPTIVE_MEMORY) return LV_SWEEP_HASH;
    if (mode == ADAPTIVE_SPEED && exec > ADAPTIVE_SWITCH) return LV_SWEEP_HASH;
    return LV_SWEEP_SORT;
}

void OptimizationPhase::run(Community &c, MODE mode) {
    lv_params p;
    lv_default_params(&p);
    p.mode = mode;
    run(c, p);
}

namespace lv {

// Vertex bound of the fast/sort sweeps' empty-chunk early return
// (phaseOptimizationSort.cu:46-60, :81-86, :186-204, :235-240).
uint32_t chunk_vertex_end(Community &c, const lv_params &p, int kind, const uint8_t *active_host);

// profile = 2 times only the degree 257-1024 bucket (bucket 6), the one the
// best-move pass runs with the device to itself.
constexpr uint32_t kProfileLightMask = 1u << 6;

static SweepArgs sweep_args(Community &c, const lv_params &p, int kind, uint32_t v_end) {
    SweepState &st = *c.state;
    SweepArgs a;
    a.g = level_view(c.graph);
    a.comm = c.communities.data();
    a.target = st.target.data();
    a.tot = c.communities_weight.data();
    a.mark = st.mark.data();
    a.bucket = st.bucket.data();
    a.stats = st.stats.data() + kGlobalSection;
    a.bstats = st.stats.data();
    a.bev = (p.profile && kind != LV_SWEEP_SORT) ? st.bev : nullptr;
    a.bev_mask = p.profile == 2 ? kProfileLightMask : ~0u;
    a.m2 = (int64_t)(2.0 * c.graph.total_weight);
    a.tw = c.graph.total_weight;
    a.v_end = v_end;
    a.kind = kind;
    a.gain = p.gain;
    a.highest = (p.tie == LV_TIE_REFERENCE && kind == LV_SWEEP_HASH);
    a.tiles = st.tiles;
    a.movers = st.movers.data();
    a.n_movers = st.n_movers.data();
    a.ct = st.ct.data();
    a.wide5 = st.bounds.max_deg[5] > kBucketMaxDeg[5];
    if (c.batch) a.bt = c.batch->view();
    return a;
}

// The tuner of the bucket bounds: at the identity partition of level 0, each
// set of kBoundSet buckets the level and times two best-move passes of the
// fast sweep (after one warm-up); the fastest set is kept for the community.
// Best-move passes only write decisions, which the sweep itself rewrites; the
// targets are set back to the partition afterwards.
static BucketBounds tune_bucket_bounds(Community &C, const lv_params &p) {
    SweepState &st = *C.state;
    hipStream_t s = C.stream();
    const uint32_t n = C.graph.n_nodes;
    float best_ms = 0;
    int best = 0;
    for (int k = 0; k < kBoundSets; ++k) {
        std::copy(kBoundSet[k], kBoundSet[k] + kNumBuckets - 1, st.bounds.max_deg);
        st.setup_level(C.graph, s);
        st.build_static_lists(s);
        SweepArgs a = sweep_args(C, p, LV_SWEEP_FAST, n);
        a.bev = nullptr;
        uint32_t *const *L = st.lists(st.cur);
        move_pass(a, st.hub_tables(), L, st.count, s, st.side_streams());
        LV_HIP(hipEventRecord(st.ev[0], s));
        move_pass(a, st.hub_tables(), L, st.count, s, st.side_streams());
        move_pass(a, st.hub_tables(), L, st.count, s, st.side_streams());
        LV_HIP(hipEventRecord(st.ev[1], s));
        LV_HIP(hipEventSynchronize(st.ev[1]));
        float ms = 0;
        LV_HIP(hipEventElapsedTime(&ms, st.ev[0], st.ev[1]));
        if (k == 0 || ms < best_ms) { best_ms = ms; best = k; }
    }
    if (n)
        LV_HIP(hipMemcpyAsync(st.target.data(), C.communities.data(), n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
    BucketBounds out;
    std::copy(kBoundSet[best], kBoundSet[best] + kNumBuckets - 1, out.max_deg);
    return out;
}

void apply_bucket_bounds(Community &C, const lv_params &p) {
    SweepState &st = *C.state;
    bool given = false;
    for (int b = 0; b < kNumBuckets - 1; ++b) given = given || p.bucket_max_deg[b] != 0;
    BucketBounds want;
    if (given) {
        if (!bounds_valid(p.bucket_max_deg)) throw Error(LV_ERR_ARG, "invalid bucket_max_deg");
        std::copy(p.bucket_max_deg, p.bucket_max_deg + kNumBuckets - 1, want.max_deg);
    } else {
        if (p.autotune && !st.bounds_tuned && C.level == 0) {
            st.tuned_bounds = tune_bucket_bounds(C, p);
            st.bounds_tuned = true;
        }
        if (st.bounds_tuned) want = st.tuned_bounds;
    }
    // the level's buckets were built with st.bounds (the tuner leaves the last
    // set it tried): rebuild them only for other bounds
    if (!std::equal(want.max_deg, want.max_deg + kNumBuckets - 1, st.bounds.max_deg)) {
        st.bounds = want;
        st.setup_level(C.graph, C.stream());
    }
}

constexpr uint32_t kTuneMinVertices = 4096;   // smaller buckets keep the default tile

void autotune_tiles(Community &c, const lv_params &p, int kind, uint32_t v_end) {
    SweepState &st = *c.state;
    if (kind == LV_SWEEP_SORT) return;
    const int f = kind == LV_SWEEP_FAST ? 1 : 0;
    if (st.tuned[f]) return;
    st.tuned[f] = true;
    hipStream_t s = c.stream();
    const SweepArgs a = sweep_args(c, p, kind, v_end);   // the sweep's own vertex bound
    uint32_t *const *L = st.lists(st.cur);
    if (kind == LV_SWEEP_HASH) pack_comm_tot(a, s);
    // best-move launches only write the decisions, the same ones each time
    for (int b = 0; b < kNumBuckets - 1; ++b) {
        const uint32_t cnt = st.count[b];
        if (cnt < kTuneMinVertices) continue;
        float best_ms = 0;
        int best = 0;
        for (int k = 0; k < kTileChoices && kTileG[b][k] != 0; ++k) {
            move_bucket(a, b, k, L[b], cnt, s);   // warm-up
            LV_HIP(hipEventRecord(st.ev[0], s));
            move_bucket(a, b, k, L[b], cnt, s);
            move_bucket(a, b, k, L[b], cnt, s);
            LV_HIP(hipEventRecord(st.ev[1], s));
            LV_HIP(hipEventSynchronize(st.ev[1]));
            float ms = 0;
            LV_HIP(hipEventElapsedTime(&ms, st.ev[0], st.ev[1]));
            if (k == 0 || ms < best_ms) { best_ms = ms; best = k; }
        }
        st.tiles.sel[b][f] = (uint8_t)best;
    }
}

// One sweep from the current lists; returns stats.  Kernels only; the caller
// syncs.  Used by the sweep loop and by lv_best_moves.
void launch_sweep(Community &c, const lv_params &p, int kind, bool apply, uint32_t v_end,
                  bool pre_decided) {
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    if (p.autotune && !pre_decided) autotune_tiles(c, p, kind, v_end);
    SweepArgs a = sweep_args(c, p, kind, v_end);
    if (apply) st.lists_static = false;   // the finalize pass writes the next lists
    sweep_reset(st.stats.data(), st.stats.size(), a.n_movers, st.next_counts(), s);
    a.counters_zeroed = true;
    uint32_t *const *cur = st.lists(st.cur);
    if (p.profile) LV_HIP(hipEventRecord(st.ev[0], s));
    if (pre_decided) {
        // the contraction's decisions and best-move statistics (one section)
        LV_HIP(hipMemcpyAsync(st.stats.data(), st.pre_stats.data(),
                              kSection * sizeof(unsigned long long), hipMemcpyDeviceToDevice, s));
        if (a.g.n)
            LV_HIP(hipMemcpyAsync(st.target.data(), st.pre_target.data(), a.g.n * sizeof(uint32_t),
                                  hipMemcpyDeviceToDevice, s));
    } else if (kind == LV_SWEEP_SORT) {
        sort_sweep(a, cur, st.count, st.bucket_off, st.ws, s);
    } else {
        move_pass(a, st.hub_tables(), cur, st.count, s,
                  p.profile == 3 ? nullptr : st.side_streams());
    }
    if (p.profile) LV_HIP(hipEventRecord(st.ev[1], s));
    if (apply) {
        commit_pass(a, cur, st.count, s, st.side_streams());
        finalize_pass(a, st.lists(st.cur ^ 1), st.next_counts(), s);
    }
    if (p.profile) LV_HIP(hipEventRecord(st.ev[2], s));
    // statistics and (after a commit) the next lists' sizes, adjacent on both sides
    LV_HIP(hipMemcpyAsync(st.host_stats, st.stats.data(),
                          kStatWordsOffset * sizeof(unsigned long long) +
                              (apply ? kNumBuckets * sizeof(uint32_t) : 0),
                          hipMemcpyDeviceToHost, s));
}

}  // namespace lv

int OptimizationPhase::run(Community &c, const lv_params &p, lv_sweep_rec *rec, int rec_cap,
                           int hash_trace, uint64_t *traversed) {
    // phaseOptimization.cuh:52: every vertex starts active (the level's setup
    // may have built these lists already)
    if (!c.state->lists_static) c.state->build_static_lists(c.stream());
    return sweep_loop(c, p, 0, c.level, rec, rec_cap, hash_trace, traversed);
}

int OptimizationPhase::sweep_loop(Community &c, const lv_params &p, int64_t exec0,
                                  int32_t rec_level, lv_sweep_rec *rec, int rec_cap,
                                  int hash_trace, uint64_t *traversed) {
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    const uint32_t n = c.graph.n_nodes;
    int64_t exec = exec0;
    double delta;
    int nrec = 0;
    std::vector<uint32_t> comm_host;
    // the first sweep of a level reached by a contraction that decided it
    // (every vertex listed, identity partition: the fast sweep it replaces)
    const bool pre = exec0 == 0 && st.pre_valid && st.pre_level == c.level &&
                     st.pre_gain == p.gain && !p.chunk_emulation &&
                     sweep_kind(p.mode, exec0) == LV_SWEEP_FAST;
    st.pre_valid = false;
    do {
        const double old_q = c.modularity;
        const int kind = sweep_kind(p.mode, exec);
        const bool pre_decided = pre && exec == exec0;
        uint32_t n_active = 0;
        for (int b = 0; b < kNumBuckets; ++b) n_active += st.count[b];
        uint32_t v_end = n;
        if (p.chunk_emulation && kind != LV_SWEEP_HASH) v_end = chunk_vertex_end(c, p, kind, nullptr);
        launch_sweep(c, p, kind, true, v_end, pre_decided);
        wait_stream(s);
        const unsigned long long *h = st.host_stats;
        const uint64_t e_active = all_bucket_slots(h, ST_E_ACTIVE);
        const uint64_t considered = all_bucket_slots(h, ST_CONSIDERED);
        const uint64_t moved = sum_slots(h, ST_MOVED);
        const int64_t dsig = (int64_t)sum_slots(h, ST_DSIGMA);
        const uint64_t sq = sum_slots(h, ST_SUMSQ);
        const uint64_t ment = sum_slots(h, ST_MOVER_ENTRIES);
        uint32_t next_counts[kNumBuckets];
        std::memcpy(next_counts, h + kStatWordsOffset, sizeof(next_counts));
        float ms_move = 0, ms_sweep = 0;
        uint64_t sweep_cand = 0, sweep_bytes = 0, commit_bytes = 0;
        if (p.profile && pre_decided) {
            // no best-move pass ran: its work is in the contraction's time
            LV_HIP(hipEventElapsedTime(&ms_sweep, st.ev[0], st.ev[2]));
            for (int b = 0; b < kNumBuckets; ++b) sweep_cand += bucket_slots(h, b, ST_CANDIDATES);
            st.acc.sweep_ms += ms_sweep;
            st.acc.sweeps += 1;
            commit_bytes = 8ull * n_active + 24ull * moved + 5ull * ment + n;
            st.acc.commit_passes += 1;
            st.acc.commit_ms += ms_sweep;
            st.acc.commit_bytes += commit_bytes;
            st.acc.commit_entries += ment;
        } else if (p.profile) {
            LV_HIP(hipEventElapsedTime(&ms_move, st.ev[0], st.ev[1]));
            LV_HIP(hipEventElapsedTime(&ms_sweep, st.ev[0], st.ev[2]));
            const bool fast = kind == LV_SWEEP_FAST;
            const bool wtd = !c.graph.unit_weights;
            // algorithmic bytes of a best-move launch, by the per-unit figures
            // of the roofline basis: per entry col + comm[x] (+ w if weighted),
            // per candidate an 8-byte total, per listed vertex row bounds,
            // comm, degree, own total and the decision (28) plus its active
            // flag (1).  The sweep adds the flags of the vertices no list holds.
            auto bytes_of = [&](uint64_t ent, uint64_t cnd, uint64_t verts, uint64_t) {
                return (wtd ? 12u : 8u) * ent + 8u * cnd + 29u * verts;
            };
            uint64_t bytes = 0;
            for (int b = 0; b < kNumBuckets; ++b) {
                const uint32_t verts = st.count[b];
                if (!verts) continue;
                const uint64_t be = bucket_slots(h, b, ST_E_ACTIVE);
                const uint64_t bc = bucket_slots(h, b, ST_CANDIDATES);
                const uint64_t bt = bucket_slots(h, b, ST_TWRITES);
                const uint64_t bb = bytes_of(be, bc, verts, bt);
                bytes += bb;
                sweep_cand += bc;
                if (kind == LV_SWEEP_SORT) continue;
                if (p.profile == 2 && !(kProfileLightMask >> b & 1u)) continue;
                float ms_b = 0;
                LV_HIP(hipEventElapsedTime(&ms_b, st.bev[2 * b], st.bev[2 * b + 1]));
                const int fk = fast ? 1 : 0;
                st.acc.bk_launches[b][fk] += 1;
                st.acc.bk_ms[b][fk] += ms_b;
                st.acc.bk_bytes[b][fk] += bb;
                st.acc.bk_entries[b][fk] += be;
                if (b == kHubBucket) continue;
                const int vi = move_variant(b, fast, wtd, st.tiles.sel[b][fast ? 1 : 0]);
                st.acc.var_launches[vi] += 1;
                st.acc.var_ms[vi] += ms_b;
                st.acc.var_bytes[vi] += bb;
                st.acc.var_entries[vi] += be;
            }
            sweep_bytes = bytes + (n - n_active);
            st.acc.move_launches += 1;
            st.acc.move_ms += ms_move;
            st.acc.move_bytes += sweep_bytes;
            st.acc.move_entries += e_active;
            st.acc.sweep_ms += ms_sweep;
            st.acc.sweeps += 1;
            // the commit by the per-unit figures of the roofline basis: per
            // listed vertex its community and target (8), per mover the two
            // total updates and its community write (24), per mover entry the
            // column id and the neighbour's flag (5), a flag per vertex (1)
            commit_bytes = 8ull * n_active + 24ull * moved + 5ull * ment + n;
            st.acc.commit_passes += 1;
            st.acc.commit_ms += ms_sweep - ms_move;
            st.acc.commit_bytes += commit_bytes;
            st.acc.commit_entries += ment;
        }
        // fast_move_update + compute_modularity (phaseOptimization.cuh:169-172)
        c.sigma_in += dsig;
        c.sum_tot_sq = sq;
        c.modularity = c.q_from_parts();
        if (traversed) *traversed += e_active;
        if (rec && nrec < rec_cap) {
            lv_sweep_rec &r = rec[nrec];
            std::memset(&r, 0, sizeof(r));
            r.level = rec_level;
            r.sweep = (int32_t)exec;
            r.kind = kind;
            r.pre_decided = pre_decided ? 1 : 0;
            r.n = n;
            r.n_active = n_active;
            r.n_moved = (uint32_t)moved;
            r.n_considered = (uint32_t)considered;
            r.E = c.graph.n_entries;
            r.e_active = e_active;
            r.sigma_in = c.sigma_in;
            r.sum_tot_sq = c.sum_tot_sq;
            r.q = c.modularity;
            r.move_ms = ms_move;
            r.sweep_ms = ms_sweep;
            r.candidates = sweep_cand;
            r.move_bytes = sweep_bytes;
            r.commit_ms = ms_sweep - ms_move;
            r.commit_bytes = commit_bytes;
            r.mover_entries = ment;
            if (hash_trace) {
                comm_host.resize(n);
                LV_HIP(hipMemcpy(comm_host.data(), c.communities.data(), n * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost));
                r.comm_hash = fnv1a_u32(comm_host.data(), n);
            }
        }
        ++nrec;
        // next sweep examines the marked vertices
        st.cur ^= 1;
        std::memcpy(st.count, next_counts, sizeof(next_counts));
        ++exec;
        delta = c.modularity - old_q;
    } while (exec <= p.max_sweeps && delta > p.sweep_threshold);
    return nrec;
}

namespace lv {

uint32_t chunk_vertex_end(Community &c, const lv_params &p, int kind, const uint8_t *active_host) {
    (void)kind;
    (void)active_host;
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    const uint32_t n = c.graph.n_nodes;
    const uint64_t E = c.graph.n_entries;
    if (E <= p.chunk_entries || n == 0) return n;   // one chunk: empty only if nothing is selected
    // chunk bounds depend on the CSR only: once per level
    const ChunkKey key{c.graph.neighboorhood_sum.data(), n, E, p.chunk_entries};
    if (!(st.chunk_key == key)) {
        std::vector<uint64_t> row((size_t)n + 1);
        LV_HIP(hipMemcpyAsync(row.data(), c.graph.neighboorhood_sum.data(),
                              ((size_t)n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        wait_stream(s);
        // limit = round + STEP; limit = limit >= E ? E : row[src(limit) - 1]
        // (phaseOptimizationSort.cu:46-60); a bound that would not advance moves
        // to the next vertex boundary instead of looping forever
        std::vector<uint32_t> cv{0};
        uint64_t round = 0;
        while (round < E) {
            uint64_t limit = round + p.chunk_entries;
            if (limit >= E) {
                limit = E;
            } else {
                const uint32_t src =
                    (uint32_t)(std::upper_bound(row.begin(), row.end(), limit) - row.begin()) - 1;
                uint64_t cand = src > 0 ? row[src - 1] : 0;
                if (cand <= round) cand = row[src] > round ? row[src] : row[src + 1];
                limit = cand;
            }
            if (limit < E)
                cv.push_back((uint32_t)(std::upper_bound(row.begin(), row.end(), limit) - row.begin()) - 1);
            round = limit;
        }
        if (cv.size() > kMaxChunks) throw Error(LV_ERR_RANGE, "too many emulated chunks");
        st.chunk_cv.assign(cv.begin(), cv.end());
        uint32_t *cv_dev = st.ws.get<uint32_t>(kChunkSlot, cv.size());
        LV_HIP(hipMemcpyAsync(cv_dev, cv.data(), cv.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        uint32_t *nsel = st.ws.get<uint32_t>(kChunkSlot + 1, n);
        nonself_counts(c.graph.neighboorhood_sum.data(), c.graph.edge_destination.data(), n, nsel, s);
        st.chunk_key = key;
    }
    const uint32_t nc = (uint32_t)st.chunk_cv.size();
    unsigned long long *csel = st.ws.get<unsigned long long>(kChunkSlot + 2, nc);
    fill_u64(csel, 0ull, nc, s);
    uint32_t *const *L = st.lists(st.cur);
    for (int b = 0; b < kNumBuckets; ++b)
        chunk_selected(L[b], st.count[b], st.ws.get<uint32_t>(kChunkSlot + 1, n),
                       st.ws.get<uint32_t>(kChunkSlot, nc), nc, csel, s);
    std::vector<unsigned long long> h(nc);
    LV_HIP(hipMemcpyAsync(h.data(), csel, nc * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    // the fast/sort sweeps return at the first chunk that selects nothing
    // (phaseOptimizationSort.cu:81-86, :235-240)
    for (uint32_t k = 0; k < nc; ++k)
        if (h[k] == 0) return st.chunk_cv[k];
    return n;
}

}  // namespace lv

// -------------------------------------------------------- AggregationPhase

// Workspace slots of the aggregation phase.
namespace lv {
enum AggSlot : size_t {
    W_MAP, W_SCAN32, W_NEWC, W_VK0, W_VK1, W_MEMBERS, W_VV1, W_HIST, W_HIST_SCAN, W_MSTART, W_P,
    W_SCAN64, W_CROW, W_CBUCKET, W_CCOUNT, W_CLISTS, W_TCOL, W_TW, W_RCOUNT, W_HSIZE, W_GINFO,
    W_GTILES, W_GITEMS, W_GP0, W_GP1, W_GPCNT, W_GPOFF, W_GDCNT, W_GDOFF, W_GBASE, W_GTCNT, W_NEWCK, W_DPS, W_DPC, W_DPN
};

static int bits_for(uint64_t x) {   // smallest b with 2^b >= x (x >= 1)
    int b = 0;
    while (b < 64 && (1ull << b) < x) ++b;
    return b;
}
}  // namespace lv

namespace lv {
// The main stream waits for the contraction's row copy (no host wait).
void join_row_copy(Community &c) {
    SweepState &st = *c.state;
    if (!st.row_copy_pending) return;
    LV_HIP(hipStreamWaitEvent(c.stream(), st.side_streams()->join[0], 0));
    st.row_copy_pending = false;
}

// community_map = inclusive_scan(tot > 0) (phaseAggregation.cuh:27-43), then
// best_communities composed with it (:45); returns the device map.
uint32_t *renumber(Community &c, uint32_t &n_new) {
    hipStream_t s = c.stream();
    Workspace &ws = c.state->ws;
    const uint32_t n = c.graph.n_nodes;
    uint32_t *map = ws.get<uint32_t>(W_MAP, n);
    renumber_flags(c.communities_weight.data(), n, map, s);
    scan_u32(map, map, n, true, ws.get<uint32_t>(W_SCAN32, scan_scratch_elems(n)), s);
    c.update_best(map);
    n_new = 0;
    if (n) LV_HIP(hipMemcpyAsync(&n_new, map + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    return map;
}
}  // namespace lv

void AggregationPhase::run(Community &c, bool aggregate, MODE mode) {
    (void)mode;  // the sort and hash contractions of the reference give the same exact graph
    uint32_t n_new = 0;
    uint32_t *map = renumber(c, n_new);
    c.n_of_best_communities = n_new;                                  // :46
    if (!aggregate) return;
    const int64_t sigma_in = c.sigma_in;
    const uint64_t sum_tot_sq = c.sum_tot_sq;
    c.state->defer_copy_join = true;
    contract(c, map, n_new);
    c.state->defer_copy_join = false;
    c.level += 1;
    if (c.state->first_sweep_gain >= 0 && !c.batch) {
        c.state->pre_valid = true;
        c.state->pre_level = c.level;
        c.state->pre_gain = c.state->first_sweep_gain;
    }
    // community.update() (:174): identity partition on the contracted graph.
    // Its modularity parts equal the ones just reached (self-loop weight of a
    // community = its internal weight, new degree = old total), so they carry.
    // The update reads the new row offsets only: it runs beside the row copy.
    c.update(sigma_in, sum_tot_sq);
    join_row_copy(c);
}

// Contraction by community hashing (phaseAggregation.cuh:51-170).
void AggregationPhase::contract(Community &c, const uint32_t *map, uint32_t n_new) {
    hipStream_t s = c.stream();
    Workspace &ws = c.state->ws;
    const uint32_t n = c.graph.n_nodes;
    const uint64_t E = c.graph.n_entries;
    const LevelView g = level_view(c.graph);
    // the contracted graph goes to the level buffer not holding the current level
    GraphDevice &ng = c.state->level_buf[c.state->next_level_buf];
    c.state->next_level_buf ^= 1;
    ng.n_nodes = n_new;
    ng.total_weight = c.graph.total_weight;                           // :172
    ng.m2 = c.graph.m2;
    ng.unit_weights = false;
    ng.neighboorhood_sum.ensure((size_t)n_new + 1);
    ng.tot_weight_per_nodes.ensure(std::max<size_t>(n_new, 1));
    // 1. members of each new community: vertices sorted by new id
    uint32_t *newc = ws.get<uint32_t>(W_NEWC, n);
    uint64_t *vk0 = ws.get<uint64_t>(W_VK0, n), *vk1 = ws.get<uint64_t>(W_VK1, n);
    uint32_t *members = ws.get<uint32_t>(W_MEMBERS, n), *vv1 = ws.get<uint32_t>(W_VV1, n);
    // the next level's first sweep decided here when the caller asks for it
    // (ModularityAlgorithms::run, batch runs)
    SweepState &st = *c.state;
    const int dgain = st.first_sweep_gain;
    unsigned long long *newck = dgain >= 0 ? ws.get<unsigned long long>(W_NEWCK, n) : nullptr;
    agg_newc(c.communities.data(), map, n, newc, vk0, members, c.communities_weight.data(), newck,
             s);
    radix_sort_pairs(vk0, members, vk1, vv1, n, bits_for(n_new),
                     ws.get<uint32_t>(W_HIST, radix_hist_elems(n)),
                     ws.get<uint32_t>(W_HIST_SCAN, scan_scratch_elems(radix_hist_elems(n))), s);
    uint32_t *mstart = ws.get<uint32_t>(W_MSTART, (size_t)n_new + 1);
    uint64_t *P = ws.get<uint64_t>(W_P, (size_t)n + 1);
    agg_members(vk0, members, g.row, n, n_new, mstart, P, s);
    scan_u64(P, P, (size_t)n + 1, false, ws.get<uint64_t>(W_SCAN64, scan_scratch_elems((size_t)n + 1)), s);
    uint64_t *crow = ws.get<uint64_t>(W_CROW, (size_t)n_new + 1);
    agg_comm_rows(P, mstart, n_new, crow, s);
    // 2. bucket communities by their number of virtual entries
    uint8_t *cb = ws.get<uint8_t>(W_CBUCKET, n_new);
    uint32_t *cnt_dev = ws.get<uint32_t>(W_CCOUNT, kNumBuckets);
    bucket_ids(crow, n_new, cb, BucketBounds(), s);   // the contraction's buckets stay the default
    bucket_counts(cb, n_new, cnt_dev, s);
    uint32_t cnt[kNumBuckets];
    LV_HIP(hipMemcpyAsync(cnt, cnt_dev, sizeof(cnt), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    uint32_t *clists = ws.get<uint32_t>(W_CLISTS, n_new);
    uint32_t *lp[kNumBuckets];
    {
        uint64_t off = 0;
        for (int b = 0; b < kNumBuckets; ++b) { lp[b] = clists + off; off += cnt[b]; }
    }
    build_lists_all(cb, n_new, lp, cnt_dev, s);
    // 3. sorted distinct rows of the LDS-table communities
    AggArgs x;
    x.g = g;
    x.newc = newc;
    x.members = members;
    x.mstart = mstart;
    x.P = P;
    x.crow = crow;
    x.tcol = ws.get<uint32_t>(W_TCOL, E);
    x.tw = ws.get<uint32_t>(W_TW, E);
    x.rcount = ws.get<uint32_t>(W_RCOUNT, n_new);
    // new degrees (= old community totals) first: the decisions read them
    new_degrees(c.communities_weight.data(), map, n, ng.tot_weight_per_nodes.data(), s);
    if (dgain >= 0) {
        x.dgain = dgain;
        x.m2 = (int64_t)(2.0 * c.graph.total_weight);   // as the sweeps take it
        x.tw_total = c.graph.total_weight;
        x.newck = newck;
        x.knew = ng.tot_weight_per_nodes.data();
        st.pre_target.ensure(std::max<size_t>(n_new, 1));
        st.pre_stats.ensure(kSection);
        x.dtarget = st.pre_target.data();
        x.dstats = st.pre_stats.data();
        x.dbt = st.first_sweep_bt;
        fill_u64(st.pre_stats.data(), 0ull, kSection, s);
    }
    // 4. giant communities: their entries in destination hash partitions,
    //    each partition summed per destination in an LDS table, on a side
    //    stream beside step 3's kernels (the two touch disjoint communities
    //    and disjoint buffers).  The giants' ids and sizes come to the host
    //    and their plan is made while step 3's tables already run.
    const uint32_t nh = cnt[kHubBucket];
    GiantArgs ga{};
    uint32_t n_parts = 0;
    const SideStreams *side = nh ? c.state->side_streams() : nullptr;
    uint64_t *gsize_dev = nullptr;
    if (side) {
        LV_HIP(hipEventRecord(side->fork, s));
        LV_HIP(hipStreamWaitEvent(side->s[0], side->fork, 0));
        gsize_dev = ws.get<uint64_t>(W_HSIZE, nh);
        gather_degrees(crow, lp[kHubBucket], nh, gsize_dev, side->s[0]);
    }
    agg_tables(x, lp, cnt, s);
    if (side) {
        hipStream_t gs = side->s[0];
        std::vector<uint32_t> giants(nh);
        std::vector<uint64_t> gsize(nh);
        LV_HIP(hipMemcpyAsync(giants.data(), lp[kHubBucket], nh * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, gs));
        LV_HIP(hipMemcpyAsync(gsize.data(), gsize_dev, nh * sizeof(uint64_t), hipMemcpyDeviceToHost,
                              gs));
        wait_stream(gs);
        // kept in the state: read by the async uploads below
        std::vector<GiantInfo> &info = c.state->giant_info;
        std::vector<uint64_t> &tiles = c.state->giant_tiles, &items = c.state->giant_items;
        info.resize(nh);
        tiles.clear();
        items.clear();
        uint64_t off = 0, mbase = 0;
        uint32_t pbase = 0, max_parts = 1;
        for (uint32_t i = 0; i < nh; ++i) {
            const uint64_t e = gsize[i];
            uint32_t lgp = 0;
            while (lgp < (uint32_t)kGiantMaxPartBits && (kGiantPartEntries << lgp) < e) ++lgp;
            const uint32_t lgt = std::max<uint32_t>(kGiantMinTileBits, lgp + 4);
            const uint32_t np = 1u << lgp;
            max_parts = std::max(max_parts, np);
            const uint64_t nt = (e + (1ull << lgt) - 1) >> lgt;
            info[i] = GiantInfo{off, mbase, giants[i], pbase, lgp, lgt};
            for (uint64_t t = 0; t < nt; ++t) tiles.push_back(((uint64_t)i << 32) | t);
            for (uint32_t q = 0; q < np; ++q) items.push_back(((uint64_t)i << 32) | q);
            off += e;
            mbase += nt * np;
            pbase += np;
        }
        if (tiles.size() > 0x7FFFFFFFull || items.size() > 0x7FFFFFFFull)
            throw Error(LV_ERR_RANGE, "giant community plan too large");
        GiantInfo *info_dev = ws.get<GiantInfo>(W_GINFO, nh);
        uint64_t *tiles_dev = ws.get<uint64_t>(W_GTILES, tiles.size());
        uint64_t *items_dev = ws.get<uint64_t>(W_GITEMS, items.size());
        LV_HIP(hipMemcpyAsync(info_dev, info.data(), nh * sizeof(GiantInfo), hipMemcpyHostToDevice, gs));
        LV_HIP(hipMemcpyAsync(tiles_dev, tiles.data(), tiles.size() * sizeof(uint64_t),
                              hipMemcpyHostToDevice, gs));
        LV_HIP(hipMemcpyAsync(items_dev, items.data(), items.size() * sizeof(uint64_t),
                              hipMemcpyHostToDevice, gs));
        ga.info = info_dev;
        ga.tiles = tiles_dev;
        ga.n_tiles = (uint32_t)tiles.size();
        ga.items = items_dev;
        ga.n_items = (uint32_t)items.size();
        ga.p0 = ws.get<unsigned long long>(W_GP0, off);
        ga.p1 = ws.get<unsigned long long>(W_GP1, off);
        ga.tcnt = ws.get<uint32_t>(W_GTCNT, tiles.size());
        ga.pcnt = ws.get<uint32_t>(W_GPCNT, pbase);
        ga.poff = ws.get<uint32_t>(W_GPOFF, pbase);
        ga.dcnt = ws.get<uint32_t>(W_GDCNT, pbase);
        ga.doff = ws.get<uint32_t>(W_GDOFF, pbase);
        ga.base = ws.get<uint32_t>(W_GBASE, mbase);
        ga.scan_tmp = ws.get<uint32_t>(W_SCAN32, scan_scratch_elems(pbase));
        ga.max_parts = max_parts;
        ga.n_giants = nh;
        if (dgain >= 0) {
            ga.dpart_s = ws.get<int64_t>(W_DPS, pbase);
            ga.dpart_c = ws.get<uint32_t>(W_DPC, pbase);
            ga.dpart_n = ws.get<uint32_t>(W_DPN, pbase);
        }
        n_parts = pbase;
        fill_u32(ga.pcnt, 0u, n_parts, gs);
        agg_giant(x, ga, n_parts, gs);
        LV_HIP(hipEventRecord(side->join[0], gs));
        LV_HIP(hipStreamWaitEvent(s, side->join[0], 0));
    }
    // 5. new offsets, then every row copied to place
    uint64_t *orow = ng.neighboorhood_sum.data();
    widen_u32(x.rcount, n_new, orow, s);
    set_u64(orow + n_new, 0ull, s);
    scan_u64(orow, orow, (size_t)n_new + 1, false,
             ws.get<uint64_t>(W_SCAN64, scan_scratch_elems((size_t)n_new + 1)), s);
    uint64_t e_new = 0;
    LV_HIP(hipMemcpyAsync(&e_new, orow + n_new, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    wait_stream(s);
    ng.edge_destination.ensure(std::max<size_t>(e_new, 1));
    ng.weights.ensure(std::max<size_t>(e_new, 1));
    // the rows are copied on a side stream: the next level's setup (degrees,
    // buckets, lists, hub tables and their host round trips) needs the new
    // offsets only and runs beside the copy when the caller defers the join
    const SideStreams *cs = c.state->side_streams();
    LV_HIP(hipEventRecord(cs->fork, s));
    LV_HIP(hipStreamWaitEvent(cs->s[0], cs->fork, 0));
    agg_copy(x, lp, cnt, orow, ng.edge_destination.data(), ng.weights.data(), cs->s[0]);
    agg_giant_copy(x, ga, n_parts, orow, ng.edge_destination.data(), ng.weights.data(), cs->s[0]);
    LV_HIP(hipEventRecord(cs->join[0], cs->s[0]));
    c.state->row_copy_pending = true;
    ng.n_entries = e_new;
    ng.n_links = e_new / 2;                                           // :97
    c.graph = GraphDevice::view_of(ng);
    if (!c.state->defer_copy_join) join_row_copy(c);
}


// ------------------------------------------------------------ RefinePhase

int RefinePhase::run(Community &c, const lv_params &p, int32_t rec_level, lv_sweep_rec *rec,
                     int rec_cap, int hash_trace, uint64_t *traversed) {
    if (c.batch) throw Error(LV_ERR_ARG, "refinement of a batch is not supported");
    hipStream_t s = c.stream();
    // the final partition on the level-0 graph (its ids are renumbered, < n)
    c.graph = GraphDevice::view_of(c.graph0);
    c.level = 0;
    const uint32_t n = c.graph.n_nodes;
    c.communities.ensure(std::max<size_t>(n, 1));
    c.communities_weight.ensure(std::max<size_t>(n, 1));
    if (n)
        LV_HIP(hipMemcpyAsync(c.communities.data(), c.best_communities.data(), n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
    c.state->setup_level(c.graph, s);
    if (n)
        LV_HIP(hipMemcpyAsync(c.state->target.data(), c.communities.data(), n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
    community_totals(c.communities.data(), c.graph.tot_weight_per_nodes.data(), n,
                     c.communities_weight.data(), s);
    c.compute_modularity();
    c.state->build_static_lists(s);
    const int ns = OptimizationPhase::sweep_loop(c, p, 1, rec_level, rec, rec_cap, hash_trace,
                                                 traversed);
    // renumbering: level-0 vertices are the original ones, so best = map[comm] - 1
    iota_u32(c.best_communities.data(), n, s);
    uint32_t n_new = 0;
    (void)renumber(c, n_new);
    c.n_of_best_communities = n_new;
    return ns;
}

// ---------------------------------------------------- ModularityAlgorithms

Community ModularityAlgorithms::Louvain(GraphHost &graph, MODE mode) {
    lv_params p;
    lv_default_params(&p);
    p.mode = mode;
    return Louvain(graph, p, nullptr, nullptr);
}

Community ModularityAlgorithms::Louvain(const GraphHost &graph, const lv_params &p, lv_result *r,
                                        hipStream_t stream) {
    Community C(graph, stream);                                       // :18
    run(C, p, r);
    return C;
}

void ModularityAlgorithms::run(Community &C, const lv_params &p, lv_result *r) {
    hipStream_t s = C.stream();
    apply_bucket_bounds(C, p);
    wait_stream(s);
    C.state->acc = KernelStatsAcc();
    const auto t0 = std::chrono::steady_clock::now();
    int nrec = 0, nlev = 0;
    uint64_t traversed = 0;
    bool cont;
    std::vector<uint32_t> best_host;
    do {                                                              // :31-67
        const double old_q = C.modularity;
        const int32_t lvl = C.level;
        const uint32_t n_lvl = C.graph.n_nodes;
        const uint64_t e_lvl = C.graph.n_entries;
        lv_sweep_rec *rec = (r && r->sweeps && nrec < r->sweeps_cap) ? r->sweeps + nrec : nullptr;
        const int cap = rec ? r->sweeps_cap - nrec : 0;
        const int ns = OptimizationPhase::run(C, p, rec, cap, r ? r->hash_trace : 0, &traversed);
        nrec += ns;
        cont = (C.modularity - old_q) > p.level_threshold;            // :52
        if (p.max_levels > 0 && lvl + 1 >= p.max_levels) cont = false;
        const double q_end = C.modularity;
        const bool wtd = !C.graph.unit_weights;
        if (p.profile) LV_HIP(hipEventRecord(C.state->ev[3], s));
        // the next level's first (fast) sweep is decided while its rows are
        // built, unless chunked sweeps are emulated
        C.state->first_sweep_gain = (cont && !p.chunk_emulation && !C.batch) ? p.gain : -1;
        AggregationPhase::run(C, cont, (MODE)p.mode);                 // :53
        C.state->first_sweep_gain = -1;
        float ms_contract = 0;
        uint64_t contract_bytes = 0, e_new = 0;
        if (cont) e_new = C.graph.n_entries;
        if (p.profile && cont) {
            LV_HIP(hipEventRecord(C.state->ev[4], s));
            LV_HIP(hipEventSynchronize(C.state->ev[4]));
            LV_HIP(hipEventElapsedTime(&ms_contract, C.state->ev[3], C.state->ev[4]));
            contract_bytes = (wtd ? 12ull : 8ull) * e_lvl + 8ull * e_new + 12ull * n_lvl +
                             12ull * C.n_of_best_communities;
            KernelStatsAcc &acc = C.state->acc;
            acc.contractions += 1;
            acc.contract_ms += ms_contract;
            acc.contract_bytes += contract_bytes;
            acc.contract_entries += e_lvl;
        }
        if (r && r->level_recs && nlev < r->levels_cap) {
            lv_level_rec &L = r->level_recs[nlev];
            std::memset(&L, 0, sizeof(L));
            L.level = lvl;
            L.sweeps = ns;
            L.n = n_lvl;
            L.n_comm = C.n_of_best_communities;
            L.E = e_lvl;
            L.q_start = old_q;
            L.q_end = q_end;
            L.contract_ms = ms_contract;
            L.E_new = e_new;
            L.contract_bytes = contract_bytes;
            if (r->hash_trace) {
                best_host.resize(C.n_original);
                LV_HIP(hipMemcpy(best_host.data(), C.best_communities.data(),
                                 C.n_original * sizeof(uint32_t), hipMemcpyDeviceToHost));
                L.best_hash = fnv1a_u32(best_host.data(), C.n_original);
            }
        }
        ++nlev;
        if (!cont) C.modularity = q_end;
    } while (cont);
    if (p.refine) {
        lv_sweep_rec *rec = (r && r->sweeps && nrec < r->sweeps_cap) ? r->sweeps + nrec : nullptr;
        const int cap = rec ? r->sweeps_cap - nrec : 0;
        nrec += RefinePhase::run(C, p, nlev, rec, cap, r ? r->hash_trace : 0, &traversed);
    }
    wait_stream(s);
    const auto t1 = std::chrono::steady_clock::now();
    g_last_run_stats = C.state->acc;
    if (r) {
        r->modularity = C.modularity;
        r->n_of_best_communities = C.n_of_best_communities;
        r->levels = nlev;
        r->total_sweeps = nrec;
        r->seconds = std::chrono::duration<double>(t1 - t0).count();
        r->traversed_entries = traversed;
        if (r->best_communities && C.n_original)
            LV_HIP(hipMemcpy(r->best_communities, C.best_communities.data(),
                             C.n_original * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
}
