/*
 * constants.h -- configuration shared by the host driver and the kernels.
 * Mirrors the reference's constants.h:1-23 (MODE, thresholds, STEP_ROUND);
 * BLOCK_SIZE is the gfx950 workgroup (4 wave64), not the reference's 128.
 */
#ifndef LV_CONSTANTS_H
#define LV_CONSTANTS_H

#include <cstdint>

/* constants.h:13 */
enum MODE { HASH = 0, SORT = 1, ADAPTIVE_SPEED = 2, ADAPTIVE_MEMORY = 3 };

/* One workgroup = 4 wavefronts of 64 lanes. */
const int BLOCK_SIZE = 256;
const int WAVE_SIZE = 64;
/* constants.h:15 -- used for both the sweep loop and the level loop */
const double MODULARITY_CONVERGED_THRESHOLD = 0.001;
/* constants.h:16 */
const int EARLY_STOP_LIMIT = 1000000000;
/* constants.h:18: edge-chunk size of the reference (chunk emulation only) */
const int STEP_ROUND = 50000000;
/* constants.h:19 */
const unsigned long long FLAG = 0xFFFFFFFFFFFFFFFFull;
/* The reference's optimisation schedule switch (phaseOptimization.cuh:39) */
const int ADAPTIVE_SWITCH = 4;

/* Empty slot marker of the LDS / global community tables. */
const uint32_t LV_EMPTY = 0xFFFFFFFFu;

#endif
