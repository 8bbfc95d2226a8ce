/*
 * phaseRefine.h -- refinement hook (reference: struct RefinePhase,
 * phaseRefine.cuh:1-3).  The reference defines it as an empty function that
 * nothing calls; run() stays that no-op, so default results match the
 * reference.  The opt-in refinement (lv_params.refine, or calling it
 * directly after the level loop): the final partition is projected onto the
 * level-0 graph and swept again from sweep index 1 (hash or sort sweeps by
 * the mode; no fast sweep, the partition is not the identity) until the sweep
 * threshold, then renumbered into best_communities.
 */
#ifndef LV_PHASE_REFINE_H
#define LV_PHASE_REFINE_H

#include "community.h"

struct RefinePhase {
    static void run() {}
    // Returns the number of sweeps; records them with level `rec_level`.
    static int run(Community &c, const lv_params &p, int32_t rec_level = 0,
                   lv_sweep_rec *rec = nullptr, int rec_cap = 0, int hash_trace = 0,
                   uint64_t *traversed = nullptr);
};

#endif
