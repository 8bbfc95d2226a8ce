/*
 * server.h -- served Louvain with continuous batching (no counterpart in the
 * reference, which runs one graph per process: main.cu:16-22).
 *
 * Jobs (one graph each) are submitted at any time.  The server advances in
 * union levels: at the start of each, queued jobs are admitted while their
 * estimated device footprint fits the HBM budget; the level's graph is the
 * disjoint union of every active job's current graph (the contracted graph of
 * a job that continues, the level-0 graph of a job just admitted), assembled
 * on the device.  One sweep loop then runs over the union with each job's own
 * 2m, convergence tests and counters (the batch engine), every job starting
 * its level with the reference's first (fast) sweep -- so a job's schedule,
 * and its result, are those of ModularityAlgorithms::Louvain on the graph
 * alone (modularityAlgorithms.h:14-81).  After the level's renumbering, jobs
 * whose level loop ended are retired (results copied to the host, memory
 * released) and the rest are contracted for the next level.  CSR and community
 * state stay on the device from admission to retirement.
 */
#ifndef LV_SERVER_H
#define LV_SERVER_H

#include <cstdint>
#include <memory>
#include <vector>

#include "community.h"

struct LouvainServer {
    LouvainServer(const lv_params &p, uint64_t hbm_budget_bytes, hipStream_t stream = nullptr);
    ~LouvainServer();
    LouvainServer(const LouvainServer &) = delete;
    LouvainServer &operator=(const LouvainServer &) = delete;

    // The graph is read at admission and must stay valid until the job is done.
    uint64_t submit(const GraphHost *g);
    // One union level (admission, sweeps, renumbering, retirement,
    // contraction).  Returns the number of jobs retired by it.
    uint32_t step();
    uint32_t active() const;
    uint32_t queued() const;
    // Ids of up to `max` jobs retired and not yet taken (oldest first).
    std::vector<uint64_t> take_finished(size_t max = SIZE_MAX);
    // Result of a retired job (false if the id is unknown or still running);
    // best (n0 entries) may be null.  release drops the stored result.
    bool result(uint64_t id, lv_job_result *r, uint32_t *best, size_t n, bool release);
    // Estimated device bytes a job holds while active.
    static uint64_t footprint(const GraphHost &g);
    lv_server_stats stats() const;

  private:
    struct Impl;
    std::unique_ptr<Impl> d_;
};

#endif
