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
    bucket.resize(std::max<size_t>(n, 1));
    mark.resize(std::max<size_t>(n, 1));
    target.resize(std::max<size_t>(n, 1));
    list_buf[0].resize(std::max<size_t>(n, 1));
    list_buf[1].resize(std::max<size_t>(n, 1));
    counts_dev.resize(2 * kNumBuckets);
    stats.resize((size_t)kNumStats * kStatSlots);
    if (!host_stats) {
        LV_HIP(hipHostMalloc((void **)&host_stats, kHostWords * sizeof(unsigned long long),
                             hipHostMallocDefault));
    }
    fill_u8(mark.data(), 0, n, s);
    // bucket ids and per-bucket sizes (static: degrees do not change within a level)
    bucket_ids(g.neighboorhood_sum.data(), n, bucket.data(), s);
    uint32_t *tmp_lists[kNumBuckets];
    for (int b = 0; b < kNumBuckets; ++b) tmp_lists[b] = list_buf[1].data();  // counting pass only
    // count per bucket by building into a scratch list region per bucket: first a
    // histogram (lists all aliased to one buffer, then rebuilt with true offsets)
    build_lists_all(bucket.data(), n, tmp_lists, counts_dev.data(), s);
    LV_HIP(hipMemcpyAsync(host_stats, counts_dev.data(), kNumBuckets * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    uint32_t cnt[kNumBuckets];
    std::memcpy(cnt, host_stats, sizeof(cnt));
    uint64_t off = 0;
    for (int b = 0; b < kNumBuckets; ++b) {
        bucket_size[b] = cnt[b];
        bucket_off[b] = off;
        off += cnt[b];
    }
    // hub tables
    n_hubs = bucket_size[kHubBucket];
    hub_slot_of.release();
    hub_tab_off.release();
    hub_tab_log2.release();
    hub_key.release();
    hub_val.release();
    hub_part_s.release();
    hub_part_c.release();
    hub_self.release();
    if (n_hubs > 0) {
        build_static_lists(s);
        std::vector<uint32_t> hubs(n_hubs);
        LV_HIP(hipMemcpyAsync(hubs.data(), list_buf[0].data() + bucket_off[kHubBucket],
                              n_hubs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        std::vector<uint64_t> rowh(g.n_nodes + 1);
        LV_HIP(hipMemcpyAsync(rowh.data(), g.neighboorhood_sum.data(),
                              (g.n_nodes + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        std::vector<uint64_t> toff(n_hubs);
        std::vector<uint32_t> tlog(n_hubs);
        uint64_t pool = 0;
        std::vector<uint32_t> slot_pairs;  // (vertex, slot)
        for (uint32_t i = 0; i < n_hubs; ++i) {
            const uint64_t d = rowh[hubs[i] + 1] - rowh[hubs[i]];
            uint32_t lg = 1;
            while ((1ull << lg) < 2 * d) ++lg;
            tlog[i] = lg;
            toff[i] = pool;
            pool += 1ull << lg;
        }
        hub_slot_of.resize(n);
        fill_u32(hub_slot_of.data(), LV_EMPTY, n, s);
        hub_tab_off.resize(n_hubs);
        hub_tab_log2.resize(n_hubs);
        hub_key.resize(pool);
        hub_val.resize(pool);
        hub_part_s.resize((size_t)n_hubs * kHubParts);
        hub_part_c.resize((size_t)n_hubs * kHubParts);
        hub_self.resize(n_hubs);
        fill_u32(hub_key.data(), LV_EMPTY, pool, s);
        fill_u32(hub_val.data(), 0u, pool, s);
        LV_HIP(hipMemcpyAsync(hub_tab_off.data(), toff.data(), n_hubs * sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(hub_tab_log2.data(), tlog.data(), n_hubs * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        std::vector<uint32_t> slots_host(n_hubs);
        for (uint32_t i = 0; i < n_hubs; ++i) slots_host[i] = i;
        // slot_of[hubs[i]] = i
        DeviceArray<uint32_t> hub_ids(n_hubs), hub_slots(n_hubs);
        LV_HIP(hipMemcpyAsync(hub_ids.data(), hubs.data(), n_hubs * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(hub_slots.data(), slots_host.data(), n_hubs * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        scatter_u32(hub_slot_of.data(), hub_ids.data(), hub_slots.data(), n_hubs, s);
        LV_HIP(hipStreamSynchronize(s));
    }
}

void SweepState::build_static_lists(hipStream_t s) {
    uint32_t *lp[kNumBuckets];
    for (int b = 0; b < kNumBuckets; ++b) lp[b] = list_buf[0].data() + bucket_off[b];
    build_lists_all(bucket.data(), n, lp, counts_dev.data(), s);
    cur = 0;
    for (int b = 0; b < kNumBuckets; ++b) count[b] = bucket_size[b];
}

uint32_t *const *SweepState::lists(int which) {
    for (int b = 0; b < kNumBuckets; ++b) list_ptr[which][b] = list_buf[which].data() + bucket_off[b];
    return list_ptr[which];
}

HubTables SweepState::hub_tables() {
    HubTables h;
    h.slot_of = hub_slot_of.data();
    h.tab_off = hub_tab_off.data();
    h.tab_log2 = hub_tab_log2.data();
    h.key = hub_key.data();
    h.val = hub_val.data();
    h.part_s = hub_part_s.data();
    h.part_c = hub_part_c.data();
    h.self = hub_self.data();
    return h;
}

SweepState::~SweepState() {
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
    for (int i = 0; i < kStatSlots; ++i) s += h[stat * kStatSlots + i];
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
    if (m2 >= 0x80000000ull) throw Error(LV_ERR_RANGE, "total weight 2m >= 2^31");
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
    LV_HIP(hipStreamSynchronize(s));
}

// --------------------------------------------------------------- Community

Community::Community(GraphHost &g, hipStream_t stream) { init(g, stream); }
Community::Community(const GraphHost &g, hipStream_t stream) { init(g, stream); }

void Community::init(const GraphHost &g, hipStream_t stream) {
    if (stream) {
        streams[0] = stream;
    } else {
        LV_HIP(hipStreamCreateWithFlags(&streams[0], hipStreamNonBlocking));
        own_stream_ = true;
    }
    streams[1] = streams[0];
    graph0 = GraphDevice(g, streams[0]);
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

void Community::reset() {
    graph = GraphDevice::view_of(graph0);
    level = 0;
    n_of_best_communities = graph.n_nodes;
    iota_u32(best_communities.data(), graph.n_nodes, stream());
    update();
}

Community::~Community() {
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
      graph0(std::move(o.graph0)),
      own_stream_(o.own_stream_) {
    streams[0] = o.streams[0];
    streams[1] = o.streams[1];
    o.streams[0] = o.streams[1] = nullptr;
    o.own_stream_ = false;
}

Community &Community::operator=(Community &&o) noexcept {
    if (this != &o) {
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

void Community::update() {
    const uint32_t n = graph.n_nodes;
    hipStream_t s = stream();
    communities.resize(std::max<size_t>(n, 1));
    communities_weight.resize(std::max<size_t>(n, 1));
    iota_u32(communities.data(), n, s);                               // community.h:58-60
    if (n)
        LV_HIP(hipMemcpyAsync(communities_weight.data(), graph.tot_weight_per_nodes.data(),
                              n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    state->setup_level(graph, s);
    iota_u32(state->target.data(), n, s);
    compute_modularity();
}

double Community::q_from_parts() const {
    const double tw = graph.total_weight;
    return (double)sigma_in / (2 * tw) - (double)sum_tot_sq / (4 * tw * tw);  // community.h:104
}

void Community::compute_modularity() {
    hipStream_t s = stream();
    fill_u64(state->stats.data(), 0ull, state->stats.size(), s);
    modularity_parts(level_view(graph), communities.data(), communities_weight.data(),
                     state->stats.data(), s);
    LV_HIP(hipMemcpyAsync(state->host_stats, state->stats.data(),
                          (size_t)kNumStats * kStatSlots * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
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

// One sweep from the current lists; returns stats.  Kernels only; the caller
// syncs.  Used by the sweep loop and by lv_best_moves.
void launch_sweep(Community &c, const lv_params &p, int kind, bool apply, uint32_t v_end) {
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    SweepArgs a;
    a.g = level_view(c.graph);
    a.comm = c.communities.data();
    a.target = st.target.data();
    a.tot = c.communities_weight.data();
    a.mark = st.mark.data();
    a.bucket = st.bucket.data();
    a.stats = st.stats.data();
    a.m2 = (int64_t)(2.0 * c.graph.total_weight);
    a.tw = c.graph.total_weight;
    a.v_end = v_end;
    a.kind = kind;
    a.gain = p.gain;
    a.highest = (p.tie == LV_TIE_REFERENCE && kind == LV_SWEEP_HASH);
    fill_u64(st.stats.data(), 0ull, st.stats.size(), s);
    uint32_t *const *cur = st.lists(st.cur);
    if (p.profile) LV_HIP(hipEventRecord(st.ev[0], s));
    move_pass(a, st.hub_tables(), cur, st.count, s);
    if (p.profile) LV_HIP(hipEventRecord(st.ev[1], s));
    if (apply) {
        commit_pass(a, cur, st.count, s);
        finalize_pass(a, st.lists(st.cur ^ 1), st.counts_dev.data() + kNumBuckets, s);
        LV_HIP(hipMemcpyAsync(st.host_stats + kStatWordsOffset,
                              st.counts_dev.data() + kNumBuckets, kNumBuckets * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
    }
    if (p.profile) LV_HIP(hipEventRecord(st.ev[2], s));
    LV_HIP(hipMemcpyAsync(st.host_stats, st.stats.data(),
                          (size_t)kNumStats * kStatSlots * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
}

}  // namespace lv

int OptimizationPhase::run(Community &c, const lv_params &p, lv_sweep_rec *rec, int rec_cap,
                           int hash_trace, uint64_t *traversed) {
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    const uint32_t n = c.graph.n_nodes;
    // phaseOptimization.cuh:52: every vertex starts active
    st.build_static_lists(s);
    int64_t exec = 0;
    double delta;
    int nrec = 0;
    std::vector<uint32_t> comm_host;
    do {
        const double old_q = c.modularity;
        const int kind = sweep_kind(p.mode, exec);
        uint32_t n_active = 0;
        for (int b = 0; b < kNumBuckets; ++b) n_active += st.count[b];
        uint32_t v_end = n;
        if (p.chunk_emulation && kind != LV_SWEEP_HASH) v_end = chunk_vertex_end(c, p, kind, nullptr);
        launch_sweep(c, p, kind, true, v_end);
        LV_HIP(hipStreamSynchronize(s));
        const unsigned long long *h = st.host_stats;
        const uint64_t e_active = sum_slots(h, ST_E_ACTIVE);
        const uint64_t considered = sum_slots(h, ST_CONSIDERED);
        const uint64_t cand = sum_slots(h, ST_CANDIDATES);
        const uint64_t moved = sum_slots(h, ST_MOVED);
        const int64_t dsig = (int64_t)sum_slots(h, ST_DSIGMA);
        const uint64_t sq = sum_slots(h, ST_SUMSQ);
        uint32_t next_counts[kNumBuckets];
        std::memcpy(next_counts, h + kStatWordsOffset, sizeof(next_counts));
        if (p.profile) {
            float ms_move = 0, ms_sweep = 0;
            LV_HIP(hipEventElapsedTime(&ms_move, st.ev[0], st.ev[1]));
            LV_HIP(hipEventElapsedTime(&ms_sweep, st.ev[0], st.ev[2]));
            const bool fast = kind == LV_SWEEP_FAST;
            const uint64_t per_entry = (fast ? 4u : 8u) + (c.graph.unit_weights ? 0u : 4u);
            const uint64_t bytes = per_entry * e_active + 4u * cand +
                                   (fast ? 28u : 32u) * (uint64_t)n_active + 4u * moved;
            st.acc.move_launches += 1;
            st.acc.move_ms += ms_move;
            st.acc.move_bytes += bytes;
            st.acc.move_entries += e_active;
            st.acc.sweep_ms += ms_sweep;
            st.acc.sweeps += 1;
        }
        // fast_move_update + compute_modularity (phaseOptimization.cuh:169-172)
        c.sigma_in += dsig;
        c.sum_tot_sq = sq;
        c.modularity = c.q_from_parts();
        if (traversed) *traversed += e_active;
        if (rec && nrec < rec_cap) {
            lv_sweep_rec &r = rec[nrec];
            std::memset(&r, 0, sizeof(r));
            r.level = c.level;
            r.sweep = (int32_t)exec;
            r.kind = kind;
            r.n = n;
            r.n_active = n_active;
            r.n_moved = (uint32_t)moved;
            r.n_considered = (uint32_t)considered;
            r.E = c.graph.n_entries;
            r.e_active = e_active;
            r.sigma_in = c.sigma_in;
            r.sum_tot_sq = c.sum_tot_sq;
            r.q = c.modularity;
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
    // Chunk bounds depend on the CSR only; the early return needs the number
    // of selected entries per chunk for the current active set, which the
    // device lists give.  Computed on the host from the active lists.
    SweepState &st = *c.state;
    hipStream_t s = c.stream();
    const uint32_t n = c.graph.n_nodes;
    const uint64_t E = c.graph.n_entries;
    if (E <= p.chunk_entries) return n;   // one chunk: empty only if nothing is selected
    std::vector<uint64_t> row(n + 1);
    std::vector<uint32_t> col(E);
    LV_HIP(hipMemcpyAsync(row.data(), c.graph.neighboorhood_sum.data(), (n + 1) * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, s));
    LV_HIP(hipMemcpyAsync(col.data(), c.graph.edge_destination.data(), E * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    std::vector<uint8_t> active(n, 0);
    std::vector<uint32_t> lst;
    for (int b = 0; b < kNumBuckets; ++b) {
        if (!st.count[b]) continue;
        lst.resize(st.count[b]);
        LV_HIP(hipMemcpyAsync(lst.data(), st.lists(st.cur)[b], st.count[b] * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        for (uint32_t v : lst) active[v] = 1;
    }
    LV_HIP(hipStreamSynchronize(s));
    uint64_t round = 0;
    uint32_t v = 0;
    while (round < E) {
        uint64_t limit = round + p.chunk_entries;
        if (limit >= E) {
            limit = E;
        } else {
            const uint32_t src = (uint32_t)(std::upper_bound(row.begin(), row.end(), limit) - row.begin()) - 1;
            uint64_t cand = src > 0 ? row[src - 1] : 0;
            if (cand <= round) cand = row[src] > round ? row[src] : row[src + 1];
            limit = cand;
        }
        uint64_t sel = 0;
        const uint32_t v0 = v;
        while (v < n && row[v] < limit) {
            if (active[v])
                for (uint64_t j = row[v]; j < row[v + 1]; ++j) sel += col[j] != v;
            ++v;
        }
        if (sel == 0) return v0;
        round = limit;
    }
    return n;
}

}  // namespace lv

// -------------------------------------------------------- AggregationPhase

void AggregationPhase::run(Community &c, bool aggregate, MODE mode) {
    (void)mode;  // sort and hash contraction give the same exact graph
    hipStream_t s = c.stream();
    const uint32_t n = c.graph.n_nodes;
    // community_map = inclusive_scan(tot > 0) (phaseAggregation.cuh:27-43)
    DeviceArray<uint32_t> map(std::max<size_t>(n, 1));
    DeviceArray<uint32_t> scratch(scan_scratch_elems(n));
    renumber_flags(c.communities_weight.data(), n, map.data(), s);
    scan_u32(map.data(), map.data(), n, true, scratch.data(), s);
    c.update_best(map.data());                                        // :45
    uint32_t n_new = 0;
    if (n) LV_HIP(hipMemcpyAsync(&n_new, map.data() + (n - 1), sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    c.n_of_best_communities = n_new;                                  // :46
    if (!aggregate) return;
    // contraction (phaseAggregation.cuh:51-105): relabel, sort, reduce
    const uint64_t E = c.graph.n_entries;
    GraphDevice ng;
    ng.n_nodes = n_new;
    ng.total_weight = c.graph.total_weight;                           // :172
    ng.m2 = c.graph.m2;
    ng.unit_weights = false;
    ng.neighboorhood_sum.resize((size_t)n_new + 1);
    ng.tot_weight_per_nodes.resize(std::max<size_t>(n_new, 1));
    uint64_t e_new = 0;
    {
        DeviceArray<uint64_t> k0(std::max<size_t>(E, 1)), k1(std::max<size_t>(E, 1));
        DeviceArray<uint32_t> v0(std::max<size_t>(E, 1)), v1(std::max<size_t>(E, 1));
        relabel_entries(level_view(c.graph), c.communities.data(), map.data(), n_new, k0.data(),
                        v0.data(), s);
        int bits = 0;
        const uint64_t maxkey = (uint64_t)n_new * (uint64_t)n_new;
        while (bits < 64 && (1ull << bits) < maxkey) ++bits;
        DeviceArray<uint32_t> hist(radix_hist_elems(E));
        DeviceArray<uint32_t> hs(scan_scratch_elems(radix_hist_elems(E)));
        radix_sort_pairs(k0.data(), v0.data(), k1.data(), v1.data(), E, bits, hist.data(), hs.data(),
                         s);
        // run heads -> inclusive scan -> 1-based output index per entry
        uint32_t *heads = v1.data();
        run_heads(k0.data(), E, heads, s);
        DeviceArray<uint32_t> es(scan_scratch_elems(E));
        scan_u32(heads, heads, E, true, es.data(), s);
        if (E) LV_HIP(hipMemcpyAsync(&e_new, heads + (E - 1), sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        uint32_t e_new32 = (uint32_t)e_new;
        (void)e_new32;
        ng.edge_destination.resize(std::max<size_t>(e_new, 1));
        ng.weights.resize(std::max<size_t>(e_new, 1));
        fill_u32(ng.weights.data(), 0u, e_new, s);
        reduce_runs(k0.data(), v0.data(), heads, E, ng.edge_destination.data(), ng.weights.data(),
                    ng.neighboorhood_sum.data(), n_new, s);
        LV_HIP(hipMemcpyAsync(ng.neighboorhood_sum.data() + n_new, &e_new, sizeof(uint64_t),
                              hipMemcpyHostToDevice, s));
        new_degrees(c.communities_weight.data(), map.data(), n, ng.tot_weight_per_nodes.data(), s);
        LV_HIP(hipStreamSynchronize(s));
    }
    ng.n_entries = e_new;
    ng.n_links = e_new / 2;                                           // :97
    c.graph = std::move(ng);
    c.level += 1;
    c.update();                                                       // :174
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
    LV_HIP(hipStreamSynchronize(s));
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
        AggregationPhase::run(C, cont, (MODE)p.mode);                 // :53
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
    LV_HIP(hipStreamSynchronize(s));
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
