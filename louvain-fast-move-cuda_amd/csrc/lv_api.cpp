// lv_api.cpp -- the C ABI (include/lv_api.h) over the C++ classes.
// Nothing throws across the boundary: every entry point maps lv::Error and
// std::bad_alloc to an lv_status.
#include "../../include/lv_api.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/batch.h"
#include "../../include/community.h"
#include "../../include/server.h"
#include "../../include/modularityAlgorithms.h"
#include "engine_state.h"
#include "lv_ingest.h"
#include "lv_kernels.h"

struct lv_graph {
    GraphHost g;
};
struct lv_batch {
    GraphBatch b;
    lv_batch(const std::vector<const GraphHost *> &g, hipStream_t s) : b(g, s) {}
};
struct lv_server {
    LouvainServer s;
    lv_server(const lv_params &p, uint64_t b, hipStream_t st) : s(p, b, st) {}
};
struct lv_community {
    Community c;
    explicit lv_community(const GraphHost &g, hipStream_t s) : c(g, s) {}
    lv_community(GraphDevice &&g, hipStream_t s) : c(std::move(g), s) {}
};

namespace {

template <typename F>
lv_status guarded(F &&f) {
    try {
        f();
        return LV_OK;
    } catch (const lv::Error &e) {
        return e.status;
    } catch (const std::bad_alloc &) {
        return LV_ERR_OOM;
    } catch (...) {
        return LV_ERR_HIP;
    }
}

lv_status need_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return LV_ERR_NO_DEVICE;
    }
    return LV_OK;
}

}  // namespace

namespace lv {
void community_totals(const uint32_t *comm, const uint32_t *kdeg, uint32_t n, uint32_t *tot,
                      hipStream_t s);
uint32_t chunk_vertex_end(Community &c, const lv_params &p, int kind, const uint8_t *active_host);
}

extern "C" {

void lv_default_params(lv_params *p) {
    p->mode = LV_HASH;
    p->gain = LV_GAIN_EXACT;
    p->tie = LV_TIE_LOWEST;
    p->chunk_emulation = 0;
    p->chunk_entries = (uint64_t)STEP_ROUND;
    p->sweep_threshold = MODULARITY_CONVERGED_THRESHOLD;
    p->level_threshold = MODULARITY_CONVERGED_THRESHOLD;
    p->max_sweeps = EARLY_STOP_LIMIT;
    p->max_levels = 0;
    p->profile = 0;
    p->autotune = 0;
    p->refine = 0;
    for (int b = 0; b < 8; ++b) p->bucket_max_deg[b] = 0;
}

const char *lv_status_string(lv_status s) {
    switch (s) {
        case LV_OK: return "ok";
        case LV_ERR_ARG: return "bad argument";
        case LV_ERR_OOM: return "out of memory";
        case LV_ERR_PARSE: return "no edge parsed";
        case LV_ERR_IO: return "cannot open file";
        case LV_ERR_NOT_DENSE: return "vertex ids are not dense";
        case LV_ERR_RANGE: return "value out of supported range";
        case LV_ERR_HIP: return "HIP runtime error";
        case LV_ERR_NO_DEVICE: return "no GPU device";
        case LV_ERR_STATE: return "call out of order";
    }
    return "unknown";
}

lv_status lv_device_count(int *count) {
    if (!count) return LV_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return LV_OK;
}

lv_status lv_graph_load_edge(const char *path, int weighted, int skip_line, lv_graph **out) {
    if (!path || !out) return LV_ERR_ARG;
    *out = nullptr;
    return guarded([&] { *out = new lv_graph{GraphHost(std::string(path), weighted != 0, skip_line)}; });
}

lv_status lv_graph_from_edges(uint64_t n_edges, const int64_t *a, const int64_t *b,
                              const uint32_t *w, lv_graph **out) {
    if (!out || (n_edges && (!a || !b))) return LV_ERR_ARG;
    *out = nullptr;
    return guarded([&] { *out = new lv_graph{GraphHost::from_edges(n_edges, a, b, w)}; });
}

lv_status lv_graph_from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                            const uint32_t *w, double total_weight, lv_graph **out) {
    if (!out || !row || (E && !col)) return LV_ERR_ARG;
    *out = nullptr;
    return guarded([&] { *out = new lv_graph{GraphHost::from_csr(n, E, row, col, w, total_weight)}; });
}

void lv_graph_free(lv_graph *g) { delete g; }

static void fill_ingest_info(lv_ingest_info *info, const GraphDevice &g, int64_t min_id,
                             const lv::IngestTimes &t) {
    if (!info) return;
    info->min_id = min_id;
    info->n = g.n_nodes;
    info->E = g.n_entries;
    info->total_weight = g.total_weight;
    info->seconds_read_upload = t.read_upload;
    info->seconds_tokenize = t.tokenize;
    info->seconds_build = t.build;
}

lv_status lv_community_load_edge(const char *path, int weighted, int skip_line, void *stream,
                                 lv_community **out, lv_ingest_info *info) {
    if (!path || !out) return LV_ERR_ARG;
    *out = nullptr;
    if (lv_status st = need_device()) return st;
    return guarded([&] {
        hipStream_t s = (hipStream_t)stream;
        hipStream_t own = nullptr;
        if (!s) {
            LV_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
            s = own;
        }
        try {
            int64_t min_id = 0;
            lv::IngestTimes t;
            GraphDevice g = lv::ingest_edge_file(path, weighted != 0, skip_line, s, &min_id, &t);
            fill_ingest_info(info, g, min_id, t);
            LV_HIP(hipStreamSynchronize(s));
            *out = new lv_community(std::move(g), (hipStream_t)stream);
        } catch (...) {
            if (own) (void)hipStreamDestroy(own);
            throw;
        }
        if (own) {
            LV_HIP(hipStreamSynchronize(own));
            (void)hipStreamDestroy(own);
        }
    });
}

lv_status lv_graph_load_edge_gpu(const char *path, int weighted, int skip_line, lv_graph **out,
                                 lv_ingest_info *info) {
    if (!path || !out) return LV_ERR_ARG;
    *out = nullptr;
    if (lv_status st = need_device()) return st;
    return guarded([&] {
        hipStream_t s = nullptr;
        LV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            int64_t min_id = 0;
            lv::IngestTimes t;
            GraphDevice g = lv::ingest_edge_file(path, weighted != 0, skip_line, s, &min_id, &t);
            fill_ingest_info(info, g, min_id, t);
            GraphHost h;
            h.n_nodes = g.n_nodes;
            h.n_links = g.n_links;
            h.total_weight = g.total_weight;
            h.min_id = min_id;
            h.unit_weights = g.unit_weights;
            h.row.resize((size_t)g.n_nodes + 1);
            h.col.resize(g.n_entries);
            LV_HIP(hipMemcpyAsync(h.row.data(), g.neighboorhood_sum.data(), h.row.size() * 8,
                                  hipMemcpyDeviceToHost, s));
            if (g.n_entries)
                LV_HIP(hipMemcpyAsync(h.col.data(), g.edge_destination.data(), g.n_entries * 4,
                                      hipMemcpyDeviceToHost, s));
            if (!g.unit_weights) {
                h.weights.resize(g.n_entries);
                LV_HIP(hipMemcpyAsync(h.weights.data(), g.weights.data(), g.n_entries * 4,
                                      hipMemcpyDeviceToHost, s));
            }
            LV_HIP(hipStreamSynchronize(s));
            *out = new lv_graph{std::move(h)};
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

lv_status lv_graph_info(const lv_graph *g, uint32_t *n, uint64_t *E, double *total_weight,
                        int64_t *min_id) {
    if (!g) return LV_ERR_ARG;
    if (n) *n = g->g.n_nodes;
    if (E) *E = g->g.n_entries();
    if (total_weight) *total_weight = g->g.total_weight;
    if (min_id) *min_id = g->g.min_id;
    return LV_OK;
}

lv_status lv_graph_copy_csr(const lv_graph *g, uint64_t *row, uint32_t *col, uint32_t *w) {
    if (!g) return LV_ERR_ARG;
    const GraphHost &h = g->g;
    if (row) std::memcpy(row, h.row.data(), h.row.size() * sizeof(uint64_t));
    if (col && !h.col.empty()) std::memcpy(col, h.col.data(), h.col.size() * sizeof(uint32_t));
    if (w) {
        if (h.unit_weights) std::fill(w, w + h.col.size(), 1u);
        else std::memcpy(w, h.weights.data(), h.weights.size() * sizeof(uint32_t));
    }
    return LV_OK;
}

lv_status lv_community_create(const lv_graph *g, void *stream, lv_community **out) {
    if (!g || !out) return LV_ERR_ARG;
    *out = nullptr;
    const lv_status d = need_device();
    if (d != LV_OK) return d;
    return guarded([&] { *out = new lv_community(g->g, (hipStream_t)stream); });
}

void lv_community_destroy(lv_community *c) { delete c; }

lv_status lv_optimization_phase(lv_community *c, const lv_params *p) {
    if (!c) return LV_ERR_ARG;
    lv_params d;
    lv_default_params(&d);
    return guarded([&] { OptimizationPhase::run(c->c, p ? *p : d); });
}

lv_status lv_aggregation_phase(lv_community *c, int aggregate, const lv_params *p) {
    if (!c) return LV_ERR_ARG;
    const MODE m = (MODE)(p ? p->mode : LV_HASH);
    return guarded([&] { AggregationPhase::run(c->c, aggregate != 0, m); });
}

lv_status lv_refine_phase(lv_community *c) {
    if (!c) return LV_ERR_ARG;
    RefinePhase::run();
    return LV_OK;
}

lv_status lv_refine(lv_community *c, const lv_params *p, lv_result *r) {
    if (!c) return LV_ERR_ARG;
    lv_params d;
    lv_default_params(&d);
    return guarded([&] {
        Community &C = c->c;
        lv_sweep_rec *rec = (r && r->sweeps) ? r->sweeps : nullptr;
        uint64_t traversed = 0;
        const int32_t lvl = C.level + 1;
        const int ns = RefinePhase::run(C, p ? *p : d, lvl, rec, rec ? r->sweeps_cap : 0,
                                        r ? r->hash_trace : 0, &traversed);
        LV_HIP(hipStreamSynchronize(C.stream()));
        if (r) {
            r->modularity = C.modularity;
            r->n_of_best_communities = C.n_of_best_communities;
            r->levels = 0;
            r->total_sweeps = ns;
            r->traversed_entries = traversed;
            if (r->best_communities && C.n_original)
                LV_HIP(hipMemcpy(r->best_communities, C.best_communities.data(),
                                 C.n_original * sizeof(uint32_t), hipMemcpyDeviceToHost));
        }
    });
}

double lv_modularity(const lv_community *c) { return c ? c->c.modularity : 0.0; }

lv_status lv_best_communities(const lv_community *c, uint32_t *out, size_t n, uint32_t *n_comm) {
    if (!c) return LV_ERR_ARG;
    if (n_comm) *n_comm = c->c.n_of_best_communities;
    if (!out) return LV_OK;
    if (n < c->c.n_original) return LV_ERR_ARG;
    return guarded([&] {
        LV_HIP(hipMemcpy(out, c->c.best_communities.data(), c->c.n_original * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    });
}

lv_status lv_louvain(const lv_graph *g, const lv_params *p, lv_result *r) {
    if (!g) return LV_ERR_ARG;
    const lv_status d = need_device();
    if (d != LV_OK) return d;
    lv_params dp;
    lv_default_params(&dp);
    return guarded([&] { (void)ModularityAlgorithms::Louvain(g->g, p ? *p : dp, r, nullptr); });
}

lv_status lv_community_reset(lv_community *c) {
    if (!c) return LV_ERR_ARG;
    return guarded([&] { c->c.reset(); });
}

lv_status lv_community_run(lv_community *c, const lv_params *p, lv_result *r) {
    if (!c) return LV_ERR_ARG;
    lv_params d;
    lv_default_params(&d);
    return guarded([&] { ModularityAlgorithms::run(c->c, p ? *p : d, r); });
}

lv_status lv_community_level_info(const lv_community *c, uint32_t *n, uint64_t *E, int32_t *level) {
    if (!c) return LV_ERR_ARG;
    if (n) *n = c->c.graph.n_nodes;
    if (E) *E = c->c.graph.n_entries;
    if (level) *level = c->c.level;
    return LV_OK;
}

lv_status lv_community_get_partition(const lv_community *c, uint32_t *comm, size_t n) {
    if (!c || !comm || n < c->c.graph.n_nodes) return LV_ERR_ARG;
    return guarded([&] {
        LV_HIP(hipMemcpy(comm, c->c.communities.data(), c->c.graph.n_nodes * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    });
}

lv_status lv_community_set_partition(lv_community *c, const uint32_t *comm, size_t n) {
    if (!c || !comm || n != c->c.graph.n_nodes) return LV_ERR_ARG;
    for (size_t i = 0; i < n; ++i)
        if (comm[i] >= n) return LV_ERR_ARG;
    return guarded([&] {
        Community &C = c->c;
        hipStream_t s = C.stream();
        LV_HIP(hipMemcpyAsync(C.communities.data(), comm, n * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(C.state->target.data(), C.communities.data(), n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
        lv::community_totals(C.communities.data(), C.graph.tot_weight_per_nodes.data(),
                             (uint32_t)n, C.communities_weight.data(), s);
        C.compute_modularity();
    });
}

lv_status lv_community_copy_graph(const lv_community *c, uint64_t *row, uint32_t *col, uint32_t *w,
                                  uint32_t *k) {
    if (!c) return LV_ERR_ARG;
    return guarded([&] {
        const GraphDevice &g = c->c.graph;
        std::vector<uint64_t> r((size_t)g.n_nodes + 1);
        std::vector<uint32_t> cc(g.n_entries), ww(g.n_entries, 1u);
        LV_HIP(hipMemcpy(r.data(), g.neighboorhood_sum.data(), r.size() * sizeof(uint64_t),
                         hipMemcpyDeviceToHost));
        if (g.n_entries)
            LV_HIP(hipMemcpy(cc.data(), g.edge_destination.data(), g.n_entries * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
        if (!g.unit_weights && g.n_entries)
            LV_HIP(hipMemcpy(ww.data(), g.weights.data(), g.n_entries * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
        // rows of contracted levels are kept in hash-table order on the device;
        // exported rows are put in destination order (the reference's CSR order)
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (uint32_t v = 0; v < g.n_nodes; ++v) {
            tmp.clear();
            for (uint64_t j = r[v]; j < r[v + 1]; ++j) tmp.emplace_back(cc[j], ww[j]);
            std::sort(tmp.begin(), tmp.end());
            for (uint64_t j = r[v]; j < r[v + 1]; ++j) {
                cc[j] = tmp[j - r[v]].first;
                ww[j] = tmp[j - r[v]].second;
            }
        }
        if (row) std::memcpy(row, r.data(), r.size() * sizeof(uint64_t));
        if (col) std::memcpy(col, cc.data(), cc.size() * sizeof(uint32_t));
        if (w) std::memcpy(w, ww.data(), ww.size() * sizeof(uint32_t));
        if (k && g.n_nodes)
            LV_HIP(hipMemcpy(k, g.tot_weight_per_nodes.data(), g.n_nodes * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
    });
}

lv_status lv_best_moves(lv_community *c, int kind, const lv_params *p, const uint8_t *active,
                        uint32_t *target, uint64_t *e_active) {
    if (!c || !target || kind < 0 || kind > 2) return LV_ERR_ARG;
    lv_params d;
    lv_default_params(&d);
    const lv_params &P = p ? *p : d;
    return guarded([&] {
        Community &C = c->c;
        lv::SweepState &st = *C.state;
        hipStream_t s = C.stream();
        const uint32_t n = C.graph.n_nodes;
        if (!active) {
            st.build_static_lists(s);
        } else {
            // marks -> lists through the finalize pass (target == comm, so no move is applied)
            LV_HIP(hipMemcpyAsync(st.mark.data(), active, n, hipMemcpyHostToDevice, s));
            lv::SweepArgs a{};
            a.g = lv::level_view(C.graph);
            a.comm = C.communities.data();
            a.target = st.target.data();
            a.tot = C.communities_weight.data();
            a.mark = st.mark.data();
            a.bucket = st.bucket.data();
            a.stats = st.stats.data() + lv::kGlobalSection;
            a.bstats = st.stats.data();
            lv::fill_u64(st.stats.data(), 0ull, st.stats.size(), s);
            lv::finalize_pass(a, st.lists(st.cur ^ 1), st.next_counts(), s);
            uint32_t cnt[lv::kNumBuckets];
            LV_HIP(hipMemcpyAsync(cnt, st.next_counts(), sizeof(cnt),
                                  hipMemcpyDeviceToHost, s));
            LV_HIP(hipStreamSynchronize(s));
            st.cur ^= 1;
            st.lists_static = false;
            std::memcpy(st.count, cnt, sizeof(cnt));
        }
        const uint32_t v_end = (P.chunk_emulation && kind != LV_SWEEP_HASH)
                                   ? lv::chunk_vertex_end(C, P, kind, nullptr)
                                   : n;
        lv::launch_sweep(C, P, kind, false, v_end);
        LV_HIP(hipStreamSynchronize(s));
        if (e_active) *e_active = lv::all_bucket_slots(st.host_stats, lv::ST_E_ACTIVE);
        LV_HIP(hipMemcpy(target, st.target.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        // restore target == comm
        LV_HIP(hipMemcpyAsync(st.target.data(), C.communities.data(), n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s));
        LV_HIP(hipStreamSynchronize(s));
    });
}

lv_status lv_community_modularity_parts(lv_community *c, int64_t *sigma_in, uint64_t *sum_tot_sq,
                                        double *q) {
    if (!c) return LV_ERR_ARG;
    return guarded([&] {
        c->c.compute_modularity();
        if (sigma_in) *sigma_in = c->c.sigma_in;
        if (sum_tot_sq) *sum_tot_sq = c->c.sum_tot_sq;
        if (q) *q = c->c.modularity;
    });
}

static void fill_stats(lv_kernel_stats *s, const lv::KernelStatsAcc &a) {
    s->move_launches = a.move_launches;
    s->move_ms = a.move_ms;
    s->move_bytes = a.move_bytes;
    s->move_entries = a.move_entries;
    s->sweep_ms = a.sweep_ms;
    s->sweeps = a.sweeps;
    for (int i = 0; i < lv::kMoveVariants; ++i) {
        s->variant_launches[i] = a.var_launches[i];
        s->variant_ms[i] = a.var_ms[i];
        s->variant_bytes[i] = a.var_bytes[i];
        s->variant_entries[i] = a.var_entries[i];
    }
    for (int b = 0; b < lv::kNumBuckets; ++b)
        for (int f = 0; f < 2; ++f) {
            s->bucket_launches[2 * b + f] = a.bk_launches[b][f];
            s->bucket_ms[2 * b + f] = a.bk_ms[b][f];
            s->bucket_bytes[2 * b + f] = a.bk_bytes[b][f];
            s->bucket_entries[2 * b + f] = a.bk_entries[b][f];
        }
    s->commit_passes = a.commit_passes;
    s->commit_ms = a.commit_ms;
    s->commit_bytes = a.commit_bytes;
    s->commit_entries = a.commit_entries;
    s->contractions = a.contractions;
    s->contract_ms = a.contract_ms;
    s->contract_bytes = a.contract_bytes;
    s->contract_entries = a.contract_entries;
}

lv_status lv_bandwidth_probe(uint64_t bytes, int reps, double *gbs) {
    if (bytes < 16 || reps < 1 || !gbs) return LV_ERR_ARG;
    const lv_status d = need_device();
    if (d != LV_OK) return d;
    return guarded([&] {
        bytes &= ~(uint64_t)15;
        lv::DeviceArray<uint8_t> a(bytes), b(bytes);
        hipStream_t s;
        LV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        LV_HIP(hipEventCreate(&e0));
        LV_HIP(hipEventCreate(&e1));
        lv::fill_u32(reinterpret_cast<uint32_t *>(a.data()), 1u, bytes / 4, s);
        lv::stream_copy(a.data(), b.data(), bytes, s);   // warm
        LV_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) lv::stream_copy(a.data(), b.data(), bytes, s);
        LV_HIP(hipEventRecord(e1, s));
        LV_HIP(hipEventSynchronize(e1));
        float ms = 0;
        LV_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;   // read + write
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}

lv_status lv_gather_probe(const lv_community *c, uint32_t width, int reps,
                          double *gathers_per_s) {
    if (!c || (width != 4 && width != 8) || reps < 1 || !gathers_per_s) return LV_ERR_ARG;
    const lv_status d = need_device();
    if (d != LV_OK) return d;
    return guarded([&] {
        const GraphDevice &g = c->c.graph0;   // the input graph (graph is the current level)
        const uint64_t E = g.n_entries;
        if (!E) throw lv::Error(LV_ERR_ARG, "graph has no entries");
        lv::DeviceArray<uint8_t> arr((size_t)g.n_nodes * width);
        lv::DeviceArray<unsigned long long> out(lv::kGatherProbeWords);
        hipStream_t s;
        LV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        hipEvent_t e0, e1;
        LV_HIP(hipEventCreate(&e0));
        LV_HIP(hipEventCreate(&e1));
        lv::fill_u32(reinterpret_cast<uint32_t *>(arr.data()), 1u, (size_t)g.n_nodes * width / 4, s);
        const uint32_t *col = g.edge_destination.data();
        lv::gather_probe(col, E, arr.data(), (int)width, out.data(), s);   // warm
        LV_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) lv::gather_probe(col, E, arr.data(), (int)width, out.data(), s);
        LV_HIP(hipEventRecord(e1, s));
        LV_HIP(hipEventSynchronize(e1));
        LV_HIP(hipGetLastError());
        float ms = 0;
        LV_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gathers_per_s = (double)E * reps / (ms * 1e-3);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    });
}

lv_status lv_debug_mark_bytes(int32_t on) {
    lv::set_mark_bytes(on != 0);
    return LV_OK;
}

lv_status lv_kernel_variant_name(int variant, int gain, char *buf, size_t len) {
    if (variant < 0 || variant >= lv::kMoveVariants || !buf || len == 0) return LV_ERR_ARG;
    const int *T = lv::kMoveTableT;
    const int sel = variant % lv::kTileChoices, rest = variant / lv::kTileChoices;
    const int b = rest / 4, fast = (rest / 2) % 2, wtd = rest % 2;
    const int G = lv::kTileG[b][sel];
    if (G == 0) return LV_ERR_ARG;
    const char *w = wtd ? "true" : "false";
    if (fast) std::snprintf(buf, len, "k_move_fast<%d, %d, %s>", G, gain, w);
    else std::snprintf(buf, len, "k_move_table<%d, %d, %d, %s, %s>", G, T[b], gain, w,
                       T[b] <= lv::kPackedMaxT ? "true" : "false");
    return LV_OK;
}

lv_status lv_community_get_tiles(const lv_community *c, uint8_t sel[16]) {
    if (!c || !sel) return LV_ERR_ARG;
    for (int b = 0; b < 8; ++b)
        for (int f = 0; f < 2; ++f) sel[2 * b + f] = c->c.state->tiles.sel[b][f];
    return LV_OK;
}

lv_status lv_community_get_buckets(const lv_community *c, uint32_t max_deg[8]) {
    if (!c || !max_deg) return LV_ERR_ARG;
    for (int b = 0; b < 8; ++b) max_deg[b] = c->c.state->bounds.max_deg[b];
    return LV_OK;
}

lv_status lv_community_set_tiles(lv_community *c, const uint8_t sel[16]) {
    if (!c || !sel) return LV_ERR_ARG;
    for (int b = 0; b < 8; ++b)
        for (int f = 0; f < 2; ++f)
            if (sel[2 * b + f] >= lv::kTileChoices || lv::kTileG[b][sel[2 * b + f]] == 0)
                return LV_ERR_ARG;
    for (int b = 0; b < 8; ++b)
        for (int f = 0; f < 2; ++f) c->c.state->tiles.sel[b][f] = sel[2 * b + f];
    c->c.state->tuned[0] = c->c.state->tuned[1] = true;
    return LV_OK;
}

lv_status lv_kernel_stats_get(const lv_community *c, lv_kernel_stats *s) {
    if (!c || !s) return LV_ERR_ARG;
    fill_stats(s, c->c.state->acc);
    return LV_OK;
}

lv_status lv_last_run_kernel_stats(lv_kernel_stats *s) {
    if (!s) return LV_ERR_ARG;
    fill_stats(s, lv::g_last_run_stats);
    return LV_OK;
}

lv_status lv_batch_create(const lv_graph *const *graphs, uint32_t n_graphs, void *stream,
                          lv_batch **out) {
    if (!out || !graphs || n_graphs == 0) return LV_ERR_ARG;
    *out = nullptr;
    for (uint32_t i = 0; i < n_graphs; ++i)
        if (!graphs[i]) return LV_ERR_ARG;
    if (lv_status st = need_device()) return st;
    return guarded([&] {
        std::vector<const GraphHost *> gs(n_graphs);
        for (uint32_t i = 0; i < n_graphs; ++i) gs[i] = &graphs[i]->g;
        *out = new lv_batch(gs, (hipStream_t)stream);
    });
}

void lv_batch_destroy(lv_batch *b) { delete b; }

lv_status lv_batch_info(const lv_batch *b, uint32_t *n_graphs, uint64_t *n_vertices,
                        uint64_t *n_entries) {
    if (!b) return LV_ERR_ARG;
    if (n_graphs) *n_graphs = b->b.n_graphs();
    if (n_vertices) *n_vertices = b->b.n_vertices();
    if (n_entries) *n_entries = b->b.n_entries();
    return LV_OK;
}

lv_status lv_batch_louvain(lv_batch *b, const lv_params *p, lv_batch_result *r) {
    if (!b) return LV_ERR_ARG;
    lv_params d;
    lv_default_params(&d);
    return guarded([&] { b->b.run(p ? *p : d, r); });
}

lv_status lv_server_create(const lv_params *p, uint64_t hbm_budget_bytes, void *stream,
                           lv_server **out) {
    if (!out || hbm_budget_bytes == 0) return LV_ERR_ARG;
    *out = nullptr;
    if (lv_status st = need_device()) return st;
    lv_params d;
    lv_default_params(&d);
    return guarded([&] { *out = new lv_server(p ? *p : d, hbm_budget_bytes, (hipStream_t)stream); });
}

void lv_server_destroy(lv_server *s) { delete s; }

lv_status lv_server_submit(lv_server *s, const lv_graph *g, uint64_t *job_id) {
    if (!s || !g || !job_id) return LV_ERR_ARG;
    return guarded([&] { *job_id = s->s.submit(&g->g); });
}

lv_status lv_server_step(lv_server *s, uint32_t *n_active, uint32_t *n_queued,
                         uint32_t *n_retired) {
    if (!s) return LV_ERR_ARG;
    return guarded([&] {
        const uint32_t r = s->s.step();
        if (n_retired) *n_retired = r;
        if (n_active) *n_active = s->s.active();
        if (n_queued) *n_queued = s->s.queued();
    });
}

lv_status lv_server_take_finished(lv_server *s, uint64_t *ids, size_t cap, size_t *count) {
    if (!s || !count || (cap && !ids)) return LV_ERR_ARG;
    return guarded([&] {
        const std::vector<uint64_t> v = s->s.take_finished(cap);   // the rest stay pending
        std::copy(v.begin(), v.end(), ids);
        *count = v.size();
    });
}

lv_status lv_server_result(lv_server *s, uint64_t job_id, lv_job_result *r, uint32_t *best,
                           size_t n, int release) {
    if (!s) return LV_ERR_ARG;
    lv_status st = LV_OK;
    lv_status g = guarded([&] {
        if (!s->s.result(job_id, r, best, n, release != 0)) st = LV_ERR_STATE;
    });
    return g != LV_OK ? g : st;
}

lv_status lv_server_footprint(const lv_graph *g, uint64_t *bytes) {
    if (!g || !bytes) return LV_ERR_ARG;
    *bytes = LouvainServer::footprint(g->g);
    return LV_OK;
}

lv_status lv_server_stats_get(const lv_server *s, lv_server_stats *out) {
    if (!s || !out) return LV_ERR_ARG;
    *out = s->s.stats();
    return LV_OK;
}

}  // extern "C"
