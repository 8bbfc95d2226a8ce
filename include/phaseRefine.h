/*
 * phaseRefine.h -- refinement hook (reference: struct RefinePhase,
 * phaseRefine.cuh:1-3).  The reference defines it as an empty function that
 * nothing calls; it stays a no-op here so results match the reference.
 */
#ifndef LV_PHASE_REFINE_H
#define LV_PHASE_REFINE_H

struct RefinePhase {
    static void run() {}
};

#endif
