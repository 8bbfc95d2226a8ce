/*
 * community.h -- partition state of the current level (reference:
 * struct Community, community.h:15-107).
 *
 * Same public members as the reference: graph, communities,
 * communities_weight, best_communities, n_of_best_communities, modularity,
 * streams, and update_best / update / compute_modularity.  Differences:
 * communities_weight holds exact integer totals (u32, 2m < 2^31) instead of
 * f64; modularity is computed from exact integer parts
 * (sigma_in, sum of tot^2) and combined in f64 as community.h:104 does.
 */
#ifndef LV_COMMUNITY_H
#define LV_COMMUNITY_H

#include <hip/hip_runtime.h>

#include <memory>

#include "constants.h"
#include "graph.h"
#include "lv_api.h"

namespace lv {
struct SweepState;
struct BatchState;
}

struct Community {
    GraphDevice graph;
    lv::DeviceArray<uint32_t> communities;          // comm[n]
    lv::DeviceArray<uint32_t> communities_weight;   // tot[n], exact
    lv::DeviceArray<uint32_t> best_communities;     // [n of level 0]

    unsigned int n_of_best_communities = 0;
    double modularity = -0.5;
    hipStream_t streams[2] = {nullptr, nullptr};    // streams[0]: every launch of this object

    // exact parts of `modularity`
    int64_t sigma_in = 0;
    uint64_t sum_tot_sq = 0;
    int32_t level = 0;
    uint32_t n_original = 0;
    std::unique_ptr<lv::SweepState> state;
    // set when this community is the disjoint union of a batch of graphs
    std::unique_ptr<lv::BatchState> batch;
    // level-0 CSR, kept resident so that reset() can start over without an upload
    GraphDevice graph0;

    explicit Community(GraphHost &g, hipStream_t stream = nullptr);  // community.h:31-46
    Community(const GraphHost &g, hipStream_t stream);
    // No graph yet: the owner installs each level's graph and calls update()
    // (the serving loop, whose level graphs are assembled from many jobs).
    explicit Community(hipStream_t stream);
    // From a level-0 graph already on the device (the GPU .edge ingest).
    Community(GraphDevice &&g0, hipStream_t stream);
    ~Community();
    Community(Community &&) noexcept;
    Community &operator=(Community &&) noexcept;

    // best[v] = map[comm[best[v]]] - 1 with a device map (community.h:48-55)
    void update_best(const unsigned int *map);
    // identity partition on the current graph, totals = degrees (community.h:57-65)
    void update();
    // the same after a contraction, with the modularity parts carried over
    void update(int64_t sigma_in_carry, uint64_t sum_tot_sq_carry);
    // from-scratch modularity of the current partition (community.h:68-106)
    void compute_modularity();

    // back to level 0: the resident level-0 CSR, identity partition, best = identity
    void reset();

    hipStream_t stream() const { return streams[0]; }
    double q_from_parts() const;

  private:
    void init(const GraphHost &g, hipStream_t stream);
    void init_resident(hipStream_t stream);
    void update_level(bool carry, int64_t sigma_in_carry, uint64_t sum_tot_sq_carry);
    bool own_stream_ = false;
};

#endif
