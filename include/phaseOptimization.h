/*
 * phaseOptimization.h -- the sweep loop of one level (reference:
 * class OptimizationPhase, phaseOptimization.cuh:15-208).
 * run() re-activates every vertex, runs the fast sweep, then hash or sort
 * sweeps per the MODE schedule (phaseOptimization.cuh:125-145) until the
 * modularity gain of a sweep is <= the threshold.
 */
#ifndef LV_PHASE_OPTIMIZATION_H
#define LV_PHASE_OPTIMIZATION_H

#include "community.h"

class OptimizationPhase {
  public:
    // phaseOptimization.cuh:184 (default parameters: reference thresholds)
    static void run(Community &c, MODE mode);
    // with explicit parameters; records one lv_sweep_rec per sweep if rec != nullptr
    static int run(Community &c, const lv_params &p, lv_sweep_rec *rec = nullptr, int rec_cap = 0,
                   int hash_trace = 0, uint64_t *traversed = nullptr);
    // The sweep loop from sweep index exec0 with the partition and active
    // lists as they stand; sweep records carry `rec_level`.
    static int sweep_loop(Community &c, const lv_params &p, int64_t exec0, int32_t rec_level,
                          lv_sweep_rec *rec, int rec_cap, int hash_trace, uint64_t *traversed);
    // Which sweep kind the schedule assigns to sweep `execution_number`.
    static int sweep_kind(int mode, int64_t execution_number);
};

#endif
