/*
 * phaseAggregation.h -- renumbering and contraction (reference:
 * struct AggregationPhase, phaseAggregation.cuh:18-190).
 * Non-empty communities (tot > 0) are renumbered in id order by an inclusive
 * scan, best_communities is composed with that map, and when `aggregate` is
 * set the graph is contracted into a CSR of the communities (parallel edges
 * summed, internal weight as one self-loop entry) and the partition reset.
 */
#ifndef LV_PHASE_AGGREGATION_H
#define LV_PHASE_AGGREGATION_H

#include "community.h"

struct AggregationPhase {
    static void run(Community &c, bool aggregate, MODE mode);
    // contract the current level with the renumbering map (device, 1-based)
    static void contract(Community &c, const uint32_t *map, uint32_t n_new);
};

#endif
