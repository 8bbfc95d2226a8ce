// server.cpp -- the serving loop: continuous batching of Louvain jobs.
//
// Each step is one union level.  Per job it performs what one iteration of
// ModularityAlgorithms::run does for a single graph (modularityAlgorithms.h:
// 31-67, with the sweep loop of phaseOptimization.cuh:158-181): jobs differ
// only in where their level starts (a just-admitted job at its level 0, a
// continuing job at its next level), and every job's level starts with the
// fast sweep, so sharing the sweep index within a union level is exact.
#include "../../include/server.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "../../include/phaseAggregation.h"
#include "../../include/phaseOptimization.h"
#include "engine_state.h"
#include "lv_kernels.h"
#include "lv_server.h"

using namespace lv;
using Clock = std::chrono::steady_clock;

namespace {

struct Job {
    uint64_t id = 0;
    const GraphHost *g = nullptr;
    uint32_t n0 = 0;
    uint64_t E0 = 0;
    uint64_t bytes = 0;
    double tw = 0;
    int64_t m2 = 0;
    int64_t sig = 0;
    uint64_t sq = 0;
    double q = 0, q_start = 0;
    int32_t levels = 0, sweeps = 0;
    bool cont = true;
    bool fresh = true;             // the next union takes its level-0 graph
    // level-0 CSR in the upload ring (row, col, weights at 256-byte aligned
    // offsets); guarded by the server's mutex
    uint64_t ring_bytes = 0;
    uint64_t ring_off = 0;
    bool uploaded = false;         // in the ring
    bool direct = false;           // too large for the ring: uploaded at assembly
    uint64_t best_off = 0;         // its best_communities (n0 local ids) in the arena
    uint32_t voff = 0, n = 0;      // vertex range in the current union
    uint32_t cvoff = 0, cn = 0;    // vertex range in the last contracted graph
    uint64_t ce0 = 0, cE = 0;      // its entries
    uint32_t ncomm = 0;
    lv_status status = LV_OK;
    Clock::time_point t_submit, t_admit, t_done;
    std::vector<uint32_t> best_host;
};

// Union levels stay below 2^31 entries (the aggregation and sort sweep count
// entries of a level in 32 bits); a larger single job is refused at submission.
constexpr uint64_t kMaxUnionEntries = 1ull << 31;

constexpr uint64_t kNoSpace = ~0ull;

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Ring layout of a job's level-0 CSR: rows (n0 u64), columns, weights.
uint64_t ring_row_bytes(const Job &j) { return align256((uint64_t)j.n0 * 8); }
uint64_t ring_col_bytes(const Job &j) { return align256(j.E0 * 4); }
uint64_t ring_size(const Job &j) {
    return ring_row_bytes(j) + ring_col_bytes(j) + (j.g->unit_weights ? 0 : ring_col_bytes(j));
}

double q_of(int64_t sigma_in, uint64_t sumsq, double tw) {
    return (double)sigma_in / (2 * tw) - (double)sumsq / (4 * tw * tw);   // community.h:104
}

double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct LouvainServer::Impl {
    lv_params p;
    uint64_t budget;
    uint64_t used = 0, next_id = 1;
    Community C;
    std::deque<std::unique_ptr<Job>> queue;
    std::vector<std::unique_ptr<Job>> active;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> done;
    std::vector<uint64_t> newly_done;
    GraphDevice ubuf;                       // the current union
    DeviceArray<uint32_t> best_arena[2];    // every active job's best_communities, compacted
    int arena = 0;
    DeviceArray<Segment> segs;
    DeviceArray<FreshSeg> fsegs, dsegs;
    DeviceArray<BestMove> moves;
    DeviceArray<JobBest> jobs;
    DeviceArray<uint32_t> voff_dev, idx_dev;
    DeviceArray<uint64_t> eoff_dev;
    uint32_t host_cap = 0;
    lv_server_stats st{};
    // Uploads of queued jobs into a FIFO ring of device memory, by a thread of
    // its own on a stream of its own, so that they overlap the level being
    // swept; assembly then copies device to device.
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job *> upq;                                   // waiting for upload
    std::deque<std::pair<uint64_t, uint64_t>> ring_allocs;   // (offset, bytes), oldest first
    DeviceArray<uint8_t> ring;
    uint64_t ring_cap = 0;
    static constexpr int kUploaders = 4;   // pageable copies scale with host threads
    hipStream_t ustream[kUploaders] = {};
    int device = 0;
    bool stop = false;
    std::thread uploader[kUploaders];

    Impl(const lv_params &pp, uint64_t b, hipStream_t s) : p(pp), budget(b), C(s) {
        C.batch = std::make_unique<BatchState>();
        LV_HIP(hipGetDevice(&device));
        // an eighth of the budget (at most 16 GiB); larger jobs are uploaded at assembly
        ring_cap = std::min<uint64_t>(std::max<uint64_t>(budget / 8, 1ull << 20), 16ull << 30);
        ring.resize(ring_cap);
        for (int t = 0; t < kUploaders; ++t) {
            LV_HIP(hipStreamCreateWithFlags(&ustream[t], hipStreamNonBlocking));
            uploader[t] = std::thread([this, t] { upload_loop(ustream[t]); });
        }
    }

    ~Impl() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (int t = 0; t < kUploaders; ++t) {
            if (uploader[t].joinable()) uploader[t].join();
            if (ustream[t]) (void)hipStreamDestroy(ustream[t]);
        }
    }

    // FIFO ring allocation (called with mu held)
    uint64_t ring_alloc(uint64_t bytes) {
        if (ring_allocs.empty()) return bytes <= ring_cap ? 0 : kNoSpace;
        const uint64_t head = ring_allocs.back().first + ring_allocs.back().second;
        const uint64_t tail = ring_allocs.front().first;
        if (ring_allocs.back().first >= tail) {   // not wrapped: [head, cap) then [0, tail)
            if (head + bytes <= ring_cap) return head;
            if (bytes <= tail) return 0;
            return kNoSpace;
        }
        return head + bytes <= tail ? head : kNoSpace;
    }

    // Each uploader takes the oldest job waiting, reserves its ring space (in
    // queue order, so the ring stays FIFO) and copies it on its own stream.
    void upload_loop(hipStream_t us) {
        (void)hipSetDevice(device);
        for (;;) {
            Job *j = nullptr;
            uint64_t off = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    if (!upq.empty()) {
                        Job *f = upq.front();
                        if (f->ring_bytes > ring_cap) {
                            f->direct = true;
                            upq.pop_front();
                            cv.notify_all();
                            continue;
                        }
                        off = ring_alloc(f->ring_bytes);
                        if (off != kNoSpace) {
                            ring_allocs.emplace_back(off, f->ring_bytes);
                            j = f;
                            upq.pop_front();
                            break;
                        }
                    }
                    cv.wait(lk);
                }
            }
            const GraphHost &h = *j->g;
            uint8_t *base = ring.data() + off;
            bool ok = hipMemcpyAsync(base, h.row.data(), (size_t)j->n0 * 8, hipMemcpyHostToDevice,
                                     us) == hipSuccess;
            if (ok && j->E0)
                ok = hipMemcpyAsync(base + ring_row_bytes(*j), h.col.data(), j->E0 * 4,
                                    hipMemcpyHostToDevice, us) == hipSuccess;
            if (ok && j->E0 && !h.unit_weights)
                ok = hipMemcpyAsync(base + ring_row_bytes(*j) + ring_col_bytes(*j), h.weights.data(),
                                    j->E0 * 4, hipMemcpyHostToDevice, us) == hipSuccess;
            ok = (hipStreamSynchronize(us) == hipSuccess) && ok;
            {
                std::lock_guard<std::mutex> lk(mu);
                j->ring_off = off;
                j->uploaded = true;
                if (!ok) j->direct = true;   // assembly uploads it itself (and frees the slot)
            }
            cv.notify_all();
        }
    }

    hipStream_t stream() { return C.stream(); }

    void ensure_slots(uint32_t A) {
        BatchState &B = *C.batch;
        if (A > host_cap) {
            const uint32_t cap = std::max<uint32_t>(A, 2 * host_cap);
            if (B.host_acc) LV_HIP(hipHostFree(B.host_acc));
            if (B.host_live) LV_HIP(hipHostFree(B.host_live));
            B.host_acc = nullptr;
            B.host_live = nullptr;
            LV_HIP(hipHostMalloc((void **)&B.host_acc, 3 * (size_t)cap * sizeof(unsigned long long),
                                 hipHostMallocDefault));
            LV_HIP(hipHostMalloc((void **)&B.host_live, cap, hipHostMallocDefault));
            host_cap = cap;
        }
        B.n_graphs = A;
        B.m2_dev.ensure(A);
        B.tw_dev.ensure(A);
        B.live_dev.ensure(A);
        B.acc.ensure(3 * (size_t)A);
    }

    void retire(std::unique_ptr<Job> j) {
        j->t_done = Clock::now();
        st.retired += 1;
        used -= j->bytes;
        newly_done.push_back(j->id);
        const uint64_t id = j->id;
        done[id] = std::move(j);
    }

    // Device work of a new job happens when the next union is assembled.
    void admit(std::unique_ptr<Job> j) {
        j->t_admit = Clock::now();
        st.admitted += 1;
        st.uploaded_bytes += j->g->row.size() * 8 + j->g->col.size() * 4 + j->g->weights.size() * 4;
        used += j->bytes;
        j->tw = j->g->total_weight;
        j->m2 = (int64_t)(2.0 * j->g->total_weight);
        active.push_back(std::move(j));
    }

    // The level graph: every active job's current graph side by side.  A
    // continuing job's contracted graph is copied on the device; a new job's
    // level-0 CSR is uploaded into place and shifted, and its degrees and
    // identity-partition parts (Community::update for one graph) are computed
    // there.  best_communities of the jobs are compacted into one arena.
    void assemble() {
        hipStream_t s = stream();
        const uint32_t A = (uint32_t)active.size();
        std::vector<Segment> sg;
        std::vector<FreshSeg> fs, fdirect;
        std::vector<uint32_t> voff(A + 1);
        uint64_t n_tot = 0, e_tot = 0, maxlen = 0, fmaxlen = 0;
        uint32_t fmax_n = 0;
        uint64_t fmax_e = 0;
        const GraphDevice &cg = C.graph;   // the last contraction
        for (uint32_t i = 0; i < A; ++i) {
            Job &j = *active[i];
            const uint32_t n = j.fresh ? j.n0 : j.cn;
            const uint64_t E = j.fresh ? j.E0 : j.cE;
            if (j.fresh) {
                FreshSeg f;
                f.dst_voff = (uint32_t)n_tot;
                f.n = n;
                f.slot = i;
                f.unit = j.g->unit_weights ? 1u : 0u;
                f.dst_eoff = e_tot;
                f.E = E;
                fs.push_back(f);
                fmax_n = std::max(fmax_n, n);
                fmax_e = std::max(fmax_e, E);
                if (j.direct) {
                    fdirect.push_back(f);
                    fmaxlen = std::max<uint64_t>(fmaxlen, std::max<uint64_t>(n, E));
                } else {   // from the upload ring
                    const uint8_t *base = ring.data() + j.ring_off;
                    Segment x;
                    x.row = reinterpret_cast<const uint64_t *>(base);
                    x.col = reinterpret_cast<const uint32_t *>(base + ring_row_bytes(j));
                    x.w = j.g->unit_weights ? nullptr
                                            : reinterpret_cast<const uint32_t *>(
                                                  base + ring_row_bytes(j) + ring_col_bytes(j));
                    x.kdeg = nullptr;   // computed by fresh_parts
                    x.src_voff = 0;
                    x.dst_voff = (uint32_t)n_tot;
                    x.n = n;
                    x.dst_eoff = e_tot;
                    x.E = E;
                    sg.push_back(x);
                    maxlen = std::max<uint64_t>(maxlen, std::max<uint64_t>(n, E));
                }
            } else {
                Segment x;
                x.row = cg.neighboorhood_sum.data();
                x.col = cg.edge_destination.data();
                x.w = cg.weights.data();
                x.kdeg = cg.tot_weight_per_nodes.data();
                x.src_voff = j.cvoff;
                x.dst_voff = (uint32_t)n_tot;
                x.n = n;
                x.dst_eoff = e_tot;
                x.E = E;
                sg.push_back(x);
                maxlen = std::max<uint64_t>(maxlen, std::max<uint64_t>(n, E));
            }
            j.voff = (uint32_t)n_tot;
            j.n = n;
            voff[i] = (uint32_t)n_tot;
            n_tot += n;
            e_tot += E;
            if (n_tot > 0x7FFFFFFFull) throw Error(LV_ERR_RANGE, "active jobs exceed 2^31-1 vertices");
        }
        voff[A] = (uint32_t)n_tot;
        GraphDevice &U = ubuf;
        U.neighboorhood_sum.ensure((size_t)n_tot + 1);
        U.edge_destination.ensure(std::max<size_t>(e_tot, 1));
        U.weights.ensure(std::max<size_t>(e_tot, 1));
        U.tot_weight_per_nodes.ensure(std::max<size_t>(n_tot, 1));
        U.n_nodes = (uint32_t)n_tot;
        U.n_entries = e_tot;
        U.n_links = e_tot / 2;
        U.total_weight = 1.0;   // per-job m lives in the batch view
        U.m2 = 0;
        U.unit_weights = false;
        if (!sg.empty()) {
            segs.ensure(sg.size());
            LV_HIP(hipMemcpyAsync(segs.data(), sg.data(), sg.size() * sizeof(Segment),
                                  hipMemcpyHostToDevice, s));
            segment_copy(segs.data(), (uint32_t)sg.size(), maxlen, U.neighboorhood_sum.data(),
                         U.edge_destination.data(), U.weights.data(), U.tot_weight_per_nodes.data(), s);
        }
        uint32_t k = 0;
        for (auto &jp : active) {
            if (!jp->fresh || !jp->direct) continue;
            const GraphHost &h = *jp->g;
            const FreshSeg &f = fdirect[k++];
            LV_HIP(hipMemcpyAsync(U.neighboorhood_sum.data() + f.dst_voff, h.row.data(),
                                  (size_t)f.n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
            if (f.E) {
                LV_HIP(hipMemcpyAsync(U.edge_destination.data() + f.dst_eoff, h.col.data(),
                                      f.E * sizeof(uint32_t), hipMemcpyHostToDevice, s));
                if (!f.unit)
                    LV_HIP(hipMemcpyAsync(U.weights.data() + f.dst_eoff, h.weights.data(),
                                          f.E * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            }
        }
        BatchState &B = *C.batch;
        ensure_slots(A);
        fill_u64(B.acc.data(), 0ull, 2 * (size_t)A, s);
        if (!fdirect.empty()) {
            dsegs.ensure(fdirect.size());
            LV_HIP(hipMemcpyAsync(dsegs.data(), fdirect.data(), fdirect.size() * sizeof(FreshSeg),
                                  hipMemcpyHostToDevice, s));
            fresh_shift(dsegs.data(), (uint32_t)fdirect.size(), fmaxlen, U.neighboorhood_sum.data(),
                        U.edge_destination.data(), U.weights.data(), s);
        }
        set_u64(U.neighboorhood_sum.data() + n_tot, e_tot, s);
        if (!fs.empty()) {
            fsegs.ensure(fs.size());
            LV_HIP(hipMemcpyAsync(fsegs.data(), fs.data(), fs.size() * sizeof(FreshSeg),
                                  hipMemcpyHostToDevice, s));
            fresh_parts(fsegs.data(), (uint32_t)fs.size(), fmax_n, fmax_e, U.neighboorhood_sum.data(),
                        U.edge_destination.data(), U.weights.data(), U.tot_weight_per_nodes.data(),
                        B.acc.data(), B.acc.data() + A, s);
        }
        B.gid_buf[0].ensure(std::max<size_t>(n_tot, 1));
        voff_dev.ensure((size_t)A + 1);
        LV_HIP(hipMemcpyAsync(voff_dev.data(), voff.data(), (A + 1) * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        fill_gid_ranges(voff_dev.data(), A, (uint32_t)n_tot, B.gid_buf[0].data(), s);
        B.gid = B.gid_buf[0].data();
        // best_communities: continuing jobs moved to the other arena, new jobs
        // start at the identity
        uint64_t b_tot = 0;
        for (auto &jp : active) b_tot += jp->n0;
        DeviceArray<uint32_t> &src = best_arena[arena], &dst = best_arena[arena ^ 1];
        dst.ensure(std::max<size_t>(b_tot, 1));
        std::vector<BestMove> mv(A);
        uint32_t max_n0 = 0;
        uint64_t off = 0;
        for (uint32_t i = 0; i < A; ++i) {
            Job &j = *active[i];
            mv[i] = BestMove{j.fresh ? nullptr : src.data() + j.best_off, dst.data() + off, j.n0};
            j.best_off = off;
            off += j.n0;
            max_n0 = std::max(max_n0, j.n0);
        }
        moves.ensure(A);
        LV_HIP(hipMemcpyAsync(moves.data(), mv.data(), A * sizeof(BestMove), hipMemcpyHostToDevice, s));
        move_best(moves.data(), A, max_n0, s);
        arena ^= 1;
        B.m2.resize(A);
        B.tw.resize(A);
        for (uint32_t i = 0; i < A; ++i) {
            B.m2[i] = active[i]->m2;
            B.tw[i] = active[i]->tw;
        }
        LV_HIP(hipMemcpyAsync(B.m2_dev.data(), B.m2.data(), A * sizeof(int64_t), hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(B.tw_dev.data(), B.tw.data(), A * sizeof(double), hipMemcpyHostToDevice, s));
        LV_HIP(hipMemcpyAsync(B.host_acc, B.acc.data(), 2 * (size_t)A * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        {   // the new jobs' ring slots are free again (oldest first, as allocated)
            std::lock_guard<std::mutex> lk(mu);
            for (auto &jp : active)
                if (jp->fresh && jp->uploaded) {
                    if (ring_allocs.empty() || ring_allocs.front().first != jp->ring_off)
                        throw Error(LV_ERR_STATE, "upload ring out of order");
                    ring_allocs.pop_front();
                }
        }
        cv.notify_all();
        for (uint32_t i = 0; i < A; ++i) {
            Job &j = *active[i];
            if (!j.fresh) continue;
            j.sig = (int64_t)B.host_acc[i];
            j.sq = B.host_acc[A + i];
            j.q = q_of(j.sig, j.sq, j.tw);
            j.fresh = false;
        }
        C.graph = GraphDevice::view_of(U);
        C.level = 0;
        C.update(0, 0);   // identity partition; per-job parts live on the host
    }
};

LouvainServer::LouvainServer(const lv_params &p, uint64_t hbm_budget_bytes, hipStream_t stream)
    : d_(std::make_unique<Impl>(p, hbm_budget_bytes, stream)) {
    if (p.chunk_emulation) throw Error(LV_ERR_ARG, "chunk emulation is not served");
    if (p.refine) throw Error(LV_ERR_ARG, "refinement is not served");
}

LouvainServer::~LouvainServer() = default;

uint64_t LouvainServer::footprint(const GraphHost &g) {
    // level graph col + w (8 B/entry), staging copy at admission (8), the
    // contracted copy (8), aggregation staging (8), radix-sort room (16);
    // per vertex: row, partition, totals, lists, graph id, best (~96)
    const uint64_t n = g.n_nodes, E = g.col.size();
    return 48 * E + 96 * n + (1ull << 20);
}

uint64_t LouvainServer::submit(const GraphHost *g) {
    Impl &d = *d_;
    if (!g || g->n_nodes == 0) throw Error(LV_ERR_ARG, "empty graph");
    if (2.0 * g->total_weight >= 2147483648.0) throw Error(LV_ERR_RANGE, "total weight 2m >= 2^31");
    uint64_t wsum = g->col.size();   // sum of the degrees, which the u32 totals hold
    if (!g->unit_weights) {
        wsum = 0;
        for (uint32_t x : g->weights) wsum += x;
    }
    if (wsum >= 0x80000000ull) throw Error(LV_ERR_RANGE, "total weight 2m >= 2^31");
    auto j = std::make_unique<Job>();
    j->id = d.next_id++;
    j->g = g;
    j->n0 = g->n_nodes;
    j->E0 = g->col.size();
    j->bytes = footprint(*g);
    if (j->bytes > d.budget) throw Error(LV_ERR_RANGE, "job larger than the HBM budget");
    if (j->E0 > kMaxUnionEntries) throw Error(LV_ERR_RANGE, "job over 2^31 entries");
    j->ring_bytes = ring_size(*j);
    j->t_submit = Clock::now();
    const uint64_t id = j->id;
    {
        std::lock_guard<std::mutex> lk(d.mu);
        d.upq.push_back(j.get());
        d.queue.push_back(std::move(j));
    }
    d.cv.notify_all();
    return id;
}

uint32_t LouvainServer::active() const { return (uint32_t)d_->active.size(); }
uint32_t LouvainServer::queued() const { return (uint32_t)d_->queue.size(); }

std::vector<uint64_t> LouvainServer::take_finished(size_t max) {
    std::vector<uint64_t> &v = d_->newly_done;
    const size_t k = std::min(max, v.size());
    std::vector<uint64_t> out(v.begin(), v.begin() + k);
    v.erase(v.begin(), v.begin() + k);
    return out;
}

bool LouvainServer::result(uint64_t id, lv_job_result *r, uint32_t *best, size_t n, bool release) {
    Impl &d = *d_;
    auto it = d.done.find(id);
    if (it == d.done.end()) return false;
    Job &j = *it->second;
    if (r) {
        r->status = j.status;
        r->modularity = j.q;
        r->n_of_best_communities = j.ncomm;
        r->n = j.n0;
        r->levels = j.levels;
        r->sweeps = j.sweeps;
        r->seconds_queued = secs(j.t_submit, j.t_admit);
        r->seconds_total = secs(j.t_submit, j.t_done);
    }
    if (best) {
        if (n < j.n0) throw Error(LV_ERR_ARG, "best buffer too small");
        if (j.status == LV_OK) std::memcpy(best, j.best_host.data(), j.n0 * sizeof(uint32_t));
    }
    if (release) d.done.erase(it);
    return true;
}

uint32_t LouvainServer::step() {
    Impl &d = *d_;
    const size_t done_before = d.newly_done.size();
    // admission by footprint (the first job of an idle server always fits:
    // submit rejects jobs larger than the budget)
    uint64_t e_active = 0;   // level-0 entries of the active jobs: bounds the union
    for (auto &j : d.active) e_active += j->E0;
    {
        std::unique_lock<std::mutex> lk(d.mu);
        while (!d.queue.empty()) {
            Job *nx = d.queue.front().get();
            if (!d.active.empty() &&
                (d.used + nx->bytes > d.budget || e_active + nx->E0 > kMaxUnionEntries))
                break;
            // a busy server admits uploaded jobs only; an idle one waits for the next
            if (!nx->uploaded && !nx->direct) {
                if (!d.active.empty()) break;
                d.cv.wait(lk, [nx] { return nx->uploaded || nx->direct; });
            }
            e_active += nx->E0;
            auto j = std::move(d.queue.front());
            d.queue.pop_front();
            d.admit(std::move(j));
        }
    }
    if (d.active.empty()) return (uint32_t)(d.newly_done.size() - done_before);
    const auto t0 = Clock::now();
    d.assemble();
    const auto t1 = Clock::now();
    d.st.steps += 1;
    d.st.seconds_assemble += secs(t0, t1);
    Community &C = d.C;
    BatchState &B = *C.batch;
    SweepState &st = *C.state;
    hipStream_t s = d.stream();
    const lv_params &p = d.p;
    const uint32_t A = (uint32_t)d.active.size();
    unsigned long long *dsig = B.host_acc, *sumsq = B.host_acc + A, *cnt = B.host_acc + 2 * A;
    // the sweep loop of every job at its own level (phaseOptimization.cuh:158-181)
    uint32_t running = A;
    for (uint32_t i = 0; i < A; ++i) {
        d.active[i]->q_start = d.active[i]->q;
        B.host_live[i] = 1;
    }
    st.build_static_lists(s);
    int64_t exec = 0;
    while (running > 0) {
        const int kind = OptimizationPhase::sweep_kind(p.mode, exec);
        LV_HIP(hipMemcpyAsync(B.live_dev.data(), B.host_live, A, hipMemcpyHostToDevice, s));
        fill_u64(B.acc.data(), 0ull, 2 * (size_t)A, s);
        launch_sweep(C, p, kind, true, C.graph.n_nodes);
        LV_HIP(hipMemcpyAsync(B.host_acc, B.acc.data(), 2 * (size_t)A * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        ++exec;
        for (uint32_t i = 0; i < A; ++i) {
            if (!B.host_live[i]) continue;
            Job &j = *d.active[i];
            j.sig += (int64_t)dsig[i];
            j.sq = sumsq[i];
            const double old = j.q;
            j.q = q_of(j.sig, j.sq, j.tw);
            ++j.sweeps;
            if (!(exec <= p.max_sweeps && j.q - old > p.sweep_threshold)) {
                B.host_live[i] = 0;
                --running;
            }
        }
        st.cur ^= 1;
        std::memcpy(st.count, st.host_stats + kStatWordsOffset, sizeof(st.count));
    }
    const auto t2 = Clock::now();
    d.st.seconds_sweeps += secs(t1, t2);
    // level test per job (modularityAlgorithms.h:52-53)
    bool any = false;
    for (auto &jp : d.active) {
        Job &j = *jp;
        ++j.levels;
        j.cont = j.q - j.q_start > p.level_threshold;
        if (p.max_levels > 0 && j.levels >= p.max_levels) j.cont = false;
        any = any || j.cont;
    }
    // renumbering (phaseAggregation.cuh:27-46) and every job's best_communities
    uint32_t n_new = 0;
    uint32_t *map = renumber(C, n_new);
    const uint32_t n = C.graph.n_nodes;
    fill_u64(B.acc.data() + 2 * (size_t)A, 0ull, A, s);
    batch_comm_counts(C.communities_weight.data(), B.gid, n, B.acc.data() + 2 * (size_t)A, s);
    LV_HIP(hipMemcpyAsync(cnt, B.acc.data() + 2 * (size_t)A, A * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    std::vector<JobBest> jb(A);
    std::vector<uint32_t> base(A + 1, 0);
    uint32_t max_n0 = 0;
    for (uint32_t i = 0; i < A; ++i) {
        Job &j = *d.active[i];
        j.ncomm = (uint32_t)cnt[i];
        base[i + 1] = base[i] + j.ncomm;
        jb[i] = JobBest{d.best_arena[d.arena].data() + j.best_off, j.n0, j.voff, base[i]};
        max_n0 = std::max(max_n0, j.n0);
    }
    d.jobs.ensure(A);
    LV_HIP(hipMemcpyAsync(d.jobs.data(), jb.data(), A * sizeof(JobBest), hipMemcpyHostToDevice, s));
    update_best_jobs(d.jobs.data(), A, max_n0, C.communities.data(), map, s);
    // retire the jobs whose level loop ended
    for (auto &jp : d.active)
        if (!jp->cont) {
            jp->best_host.resize(jp->n0);
            LV_HIP(hipMemcpyAsync(jp->best_host.data(), d.best_arena[d.arena].data() + jp->best_off,
                                  jp->n0 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        }
    LV_HIP(hipStreamSynchronize(s));
    std::vector<std::unique_ptr<Job>> next;
    for (uint32_t i = 0; i < A; ++i) {
        auto &jp = d.active[i];
        if (jp->cont) {
            jp->cvoff = base[i];
            jp->cn = jp->ncomm;
            next.push_back(std::move(jp));
        } else {
            d.retire(std::move(jp));
        }
    }
    d.active.swap(next);
    // contraction of the level (jobs retired above are contracted too and
    // dropped by the next assembly); entry ranges of the continuing jobs
    if (any) {
        AggregationPhase::contract(C, map, n_new);
        const uint32_t K = (uint32_t)d.active.size();
        std::vector<uint32_t> idx(2 * K);
        for (uint32_t i = 0; i < K; ++i) {
            idx[2 * i] = d.active[i]->cvoff;
            idx[2 * i + 1] = d.active[i]->cvoff + d.active[i]->cn;
        }
        d.idx_dev.ensure(2 * K);
        d.eoff_dev.ensure(2 * K);
        LV_HIP(hipMemcpyAsync(d.idx_dev.data(), idx.data(), 2 * K * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
        gather_u64(C.graph.neighboorhood_sum.data(), d.idx_dev.data(), 2 * K, d.eoff_dev.data(), s);
        std::vector<uint64_t> eo(2 * K);
        LV_HIP(hipMemcpyAsync(eo.data(), d.eoff_dev.data(), 2 * K * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < K; ++i) {
            d.active[i]->ce0 = eo[2 * i];
            d.active[i]->cE = eo[2 * i + 1] - eo[2 * i];
        }
    }
    d.st.seconds_aggregate += secs(t2, Clock::now());
    return (uint32_t)(d.newly_done.size() - done_before);
}

lv_server_stats LouvainServer::stats() const { return d_->st; }
