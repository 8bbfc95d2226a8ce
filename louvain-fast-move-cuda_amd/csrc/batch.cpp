// batch.cpp -- many graphs as one disjoint union through the level loop.
//
// Per graph this is ModularityAlgorithms::run (modularityAlgorithms.h:14-81)
// with OptimizationPhase::run's sweep loop (phaseOptimization.cuh:158-181):
// every graph keeps its own modularity, sweep and level tests and counters.
// The kernels are the single-graph ones; the graph id of a vertex selects its
// 2m and m and masks graphs that have stopped (BatchView).  Per sweep the host
// reads 2 words per graph (sigma_in change, sum of tot^2), decides which
// graphs go on, and uploads one byte per graph.
#include "../../include/batch.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

#include "../../include/phaseAggregation.h"
#include "../../include/phaseOptimization.h"
#include "engine_state.h"
#include "lv_kernels.h"

namespace lv {

BatchState::~BatchState() {
    if (host_acc) (void)hipHostFree(host_acc);
    if (host_live) (void)hipHostFree(host_live);
}

namespace {

// One GraphHost holding the graphs side by side (ids shifted by voff[g]).
GraphHost union_of(const std::vector<const GraphHost *> &gs, std::vector<uint32_t> &voff) {
    GraphHost u;
    u.disjoint_union = true;
    const size_t G = gs.size();
    voff.assign(G + 1, 0);
    std::vector<uint64_t> eoff(G + 1, 0);
    bool unit = true;
    for (size_t g = 0; g < G; ++g) {
        const uint64_t n = gs[g]->n_nodes;
        if ((uint64_t)voff[g] + n > 0x7FFFFFFFull) throw Error(LV_ERR_RANGE, "batch over 2^31-1 vertices");
        voff[g + 1] = voff[g] + (uint32_t)n;
        eoff[g + 1] = eoff[g] + gs[g]->col.size();
        unit = unit && gs[g]->unit_weights;
        u.n_links += gs[g]->n_links;
        u.total_weight += gs[g]->total_weight;
    }
    u.n_nodes = voff[G];
    u.unit_weights = unit;
    u.row.resize((size_t)u.n_nodes + 1);
    u.col.resize(eoff[G]);
    if (!unit) u.weights.resize(eoff[G]);
    // graphs are independent slices: fill them in parallel
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (size_t g = t; g < G; g += nt) {
                const GraphHost &h = *gs[g];
                const uint32_t vo = voff[g];
                const uint64_t eo = eoff[g];
                for (uint32_t v = 0; v < h.n_nodes; ++v) u.row[vo + v] = eo + h.row[v];
                for (size_t j = 0; j < h.col.size(); ++j) u.col[eo + j] = h.col[j] + vo;
                if (!unit) {
                    if (h.unit_weights) std::fill(u.weights.begin() + eo, u.weights.begin() + eo + h.col.size(), 1u);
                    else std::copy(h.weights.begin(), h.weights.end(), u.weights.begin() + eo);
                }
            }
        });
    for (auto &x : th) x.join();
    u.row[u.n_nodes] = eoff[G];
    return u;
}

double q_of(int64_t sigma_in, uint64_t sumsq, double tw) {
    return (double)sigma_in / (2 * tw) - (double)sumsq / (4 * tw * tw);   // community.h:104
}

}  // namespace
}  // namespace lv

using namespace lv;

GraphBatch::GraphBatch(const std::vector<const GraphHost *> &graphs, hipStream_t stream) {
    if (graphs.empty()) throw Error(LV_ERR_ARG, "empty batch");
    const uint32_t G = (uint32_t)graphs.size();
    auto B = std::make_unique<BatchState>();
    B->n_graphs = G;
    B->entries.resize(G);
    B->tw.resize(G);
    B->m2.resize(G);
    B->sigma0.resize(G);
    B->sumsq0.resize(G);
    // per-graph identity-partition parts and limits (as Community::update does
    // for one graph: sigma_in = self-loop weight, sum of k^2)
    for (uint32_t g = 0; g < G; ++g) {
        const GraphHost &h = *graphs[g];
        int64_t sig = 0;
        uint64_t sq = 0, m2w = 0;
        for (uint32_t v = 0; v < h.n_nodes; ++v) {
            uint64_t k = 0;
            for (uint64_t j = h.row[v]; j < h.row[v + 1]; ++j) {
                const uint32_t w = h.unit_weights ? 1u : h.weights[j];
                k += w;
                if (h.col[j] == v) sig += w;
            }
            sq += k * k;
            m2w += k;
        }
        if (m2w >= 0x80000000ull) throw Error(LV_ERR_RANGE, "a graph of the batch has 2m >= 2^31");
        B->entries[g] = h.col.size();
        B->tw[g] = h.total_weight;
        B->m2[g] = (int64_t)(2.0 * h.total_weight);
        B->sigma0[g] = sig;
        B->sumsq0[g] = sq;
    }
    std::vector<uint32_t> voff;
    {
        GraphHost u = union_of(graphs, voff);
        C_ = std::make_unique<Community>(u, stream);
    }
    B->voff = voff;
    hipStream_t s = C_->stream();
    const uint32_t n0 = C_->graph.n_nodes;
    std::vector<uint32_t> gid(n0);
    for (uint32_t g = 0; g < G; ++g) std::fill(gid.begin() + voff[g], gid.begin() + voff[g + 1], g);
    B->gid0.resize(std::max<size_t>(n0, 1));
    B->m2_dev.resize(G);
    B->tw_dev.resize(G);
    B->live_dev.resize(G);
    B->acc.resize(3 * (size_t)G);
    LV_HIP(hipHostMalloc((void **)&B->host_acc, 3 * (size_t)G * sizeof(unsigned long long),
                         hipHostMallocDefault));
    LV_HIP(hipHostMalloc((void **)&B->host_live, G, hipHostMallocDefault));
    LV_HIP(hipMemcpyAsync(B->gid0.data(), gid.data(), n0 * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    LV_HIP(hipMemcpyAsync(B->m2_dev.data(), B->m2.data(), G * sizeof(int64_t), hipMemcpyHostToDevice, s));
    LV_HIP(hipMemcpyAsync(B->tw_dev.data(), B->tw.data(), G * sizeof(double), hipMemcpyHostToDevice, s));
    wait_stream(s);
    B->gid = B->gid0.data();
    C_->batch = std::move(B);
}

GraphBatch::~GraphBatch() = default;

uint32_t GraphBatch::n_graphs() const { return C_->batch->n_graphs; }
uint64_t GraphBatch::n_vertices() const { return C_->graph0.n_nodes; }
uint64_t GraphBatch::n_entries() const { return C_->graph0.n_entries; }

void GraphBatch::run(const lv_params &p, lv_batch_result *r) {
    Community &C = *C_;
    BatchState &B = *C.batch;
    SweepState &st = *C.state;
    const uint32_t G = B.n_graphs;
    hipStream_t s = C.stream();
    if (p.refine) throw Error(LV_ERR_ARG, "refinement is not batched");
    if (p.chunk_emulation)
        for (uint32_t g = 0; g < G; ++g)
            if (B.entries[g] > p.chunk_entries)
                throw Error(LV_ERR_ARG, "chunk emulation in a batch needs graphs of one chunk");
    wait_stream(s);
    const auto t0 = std::chrono::steady_clock::now();
    C.reset();
    B.gid = B.gid0.data();
    B.next_gid = 0;
    // bucket bounds (and their tuning, with every graph sweeping)
    std::fill(B.host_live, B.host_live + G, (uint8_t)1);
    LV_HIP(hipMemcpyAsync(B.live_dev.data(), B.host_live, G, hipMemcpyHostToDevice, s));
    apply_bucket_bounds(C, p);
    std::vector<int64_t> sig(B.sigma0);
    std::vector<uint64_t> sq(B.sumsq0);
    std::vector<double> q(G), q_start(G);
    std::vector<uint8_t> done(G, 0);
    std::vector<int32_t> levels(G, 0), sweeps(G, 0);
    std::vector<uint32_t> ncomm(G, 0);
    for (uint32_t g = 0; g < G; ++g) q[g] = q_of(sig[g], sq[g], B.tw[g]);
    uint64_t traversed = 0;
    int32_t union_levels = 0, union_sweeps = 0;
    unsigned long long *dsig = B.host_acc, *sumsq = B.host_acc + G, *cnt = B.host_acc + 2 * G;
    for (;;) {                                                         // modularityAlgorithms.h:31-67
        uint32_t running = 0;
        for (uint32_t g = 0; g < G; ++g) {
            q_start[g] = q[g];
            B.host_live[g] = done[g] ? 0 : 1;
            running += B.host_live[g];
        }
        // phaseOptimization.cuh:52: every vertex starts active
        st.build_static_lists(s);
        int64_t exec = 0;
        while (running > 0) {                                          // phaseOptimization.cuh:158-181
            const int kind = OptimizationPhase::sweep_kind(p.mode, exec);
            LV_HIP(hipMemcpyAsync(B.live_dev.data(), B.host_live, G, hipMemcpyHostToDevice, s));
            fill_u64(B.acc.data(), 0ull, 2 * (size_t)G, s);
            // a level's first sweep decided by the contraction that built it
            const bool pre = exec == 0 && st.pre_valid && kind == LV_SWEEP_FAST;
            st.pre_valid = false;
            launch_sweep(C, p, kind, true, C.graph.n_nodes, pre);
            LV_HIP(hipMemcpyAsync(B.host_acc, B.acc.data(), 2 * (size_t)G * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
            wait_stream(s);
            const unsigned long long *h = st.host_stats;
            traversed += all_bucket_slots(h, ST_E_ACTIVE);
            ++exec;
            for (uint32_t g = 0; g < G; ++g) {
                if (!B.host_live[g]) continue;
                sig[g] += (int64_t)dsig[g];
                sq[g] = sumsq[g];
                const double old = q[g];
                q[g] = q_of(sig[g], sq[g], B.tw[g]);
                ++sweeps[g];
                if (!(exec <= p.max_sweeps && q[g] - old > p.sweep_threshold)) {
                    B.host_live[g] = 0;
                    --running;
                }
            }
            ++union_sweeps;
            st.cur ^= 1;
            std::memcpy(st.count, h + kStatWordsOffset, sizeof(st.count));
        }
        ++union_levels;
        // the level test per graph (modularityAlgorithms.h:52-53)
        bool any = false;
        for (uint32_t g = 0; g < G; ++g) {
            if (done[g]) continue;
            ++levels[g];
            bool cont = q[g] - q_start[g] > p.level_threshold;
            if (p.max_levels > 0 && levels[g] >= p.max_levels) cont = false;
            if (!cont) done[g] = 1;
            any = any || cont;
        }
        // renumbering for every graph (stopped graphs keep their ids: their
        // communities are single vertices that never move)
        uint32_t n_new = 0;
        uint32_t *map = renumber(C, n_new);
        C.n_of_best_communities = n_new;
        const uint32_t n = C.graph.n_nodes;
        fill_u64(B.acc.data() + 2 * (size_t)G, 0ull, G, s);
        batch_comm_counts(C.communities_weight.data(), B.gid, n, B.acc.data() + 2 * (size_t)G, s);
        if (!any) break;
        DeviceArray<uint32_t> &gn = B.gid_buf[B.next_gid];
        gn.ensure(std::max<size_t>(n_new, 1));
        batch_next_gid(C.communities_weight.data(), map, B.gid, n, gn.data(), s);
        // the contraction decides the next level's first sweep for the graphs
        // that go on (their live flags, as that sweep will upload them)
        const bool decide = !p.chunk_emulation;
        if (decide) {
            for (uint32_t g = 0; g < G; ++g) B.host_live[g] = done[g] ? 0 : 1;
            LV_HIP(hipMemcpyAsync(B.live_dev.data(), B.host_live, G, hipMemcpyHostToDevice, s));
            BatchView v = B.view();
            v.gid = gn.data();
            st.first_sweep_bt = v;
            st.first_sweep_gain = p.gain;
        }
        AggregationPhase::contract(C, map, n_new);
        st.first_sweep_gain = -1;
        st.first_sweep_bt = BatchView();
        st.pre_valid = decide;
        C.level += 1;
        B.gid = gn.data();
        B.next_gid ^= 1;
        C.update(0, 0);   // identity partition; per-graph parts carry on the host
    }
    LV_HIP(hipMemcpyAsync(cnt, B.acc.data() + 2 * (size_t)G, G * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    wait_stream(s);
    for (uint32_t g = 0; g < G; ++g) ncomm[g] = (uint32_t)cnt[g];
    const auto t1 = std::chrono::steady_clock::now();
    if (!r) return;
    r->union_levels = union_levels;
    r->union_sweeps = union_sweeps;
    r->seconds = std::chrono::duration<double>(t1 - t0).count();
    r->traversed_entries = traversed;
    for (uint32_t g = 0; g < G; ++g) {
        if (r->modularity) r->modularity[g] = q[g];
        if (r->n_of_best_communities) r->n_of_best_communities[g] = ncomm[g];
        if (r->levels) r->levels[g] = levels[g];
        if (r->sweeps) r->sweeps[g] = sweeps[g];
    }
    if (r->best_communities) {
        const uint32_t n0 = C.n_original;
        if (n0)
            LV_HIP(hipMemcpy(r->best_communities, C.best_communities.data(), n0 * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
        // the final renumbering numbers the graphs' communities one graph after
        // the other: subtract each graph's first id
        uint32_t base = 0;
        for (uint32_t g = 0; g < G; ++g) {
            for (uint32_t v = B.voff[g]; v < B.voff[g + 1]; ++v) r->best_communities[v] -= base;
            base += ncomm[g];
        }
    }
}
