/*
 * modularityAlgorithms.h -- the level loop (reference:
 * struct ModularityAlgorithms, modularityAlgorithms.h:7-82).
 * Louvain(graph, mode): optimisation phase; continue while the level's
 * modularity gain exceeds the threshold; aggregation phase (contracting only
 * when continuing).  The returned Community holds best_communities,
 * n_of_best_communities and modularity.
 */
#ifndef LV_MODULARITY_ALGORITHMS_H
#define LV_MODULARITY_ALGORITHMS_H

#include "community.h"
#include "phaseAggregation.h"
#include "phaseOptimization.h"
#include "phaseRefine.h"

struct ModularityAlgorithms {
    static Community Louvain(GraphHost &graph, MODE mode);
    static Community Louvain(const GraphHost &graph, const lv_params &p, lv_result *r = nullptr,
                             hipStream_t stream = nullptr);
    // The level loop on an existing device-resident Community (from its current level).
    static void run(Community &C, const lv_params &p, lv_result *r = nullptr);
};

#endif
