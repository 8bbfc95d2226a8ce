/*
 * operators.h -- selection rules shared by the kernels.
 * Reference: operators.h:9-70 (CountNotZero, GetMaxValue, PairSum, Square).
 */
#ifndef LV_OPERATORS_H
#define LV_OPERATORS_H

#include <hip/hip_runtime.h>
#include <cstdint>

#include "constants.h"
#include "operatorsCommunity.h"
#include "operatorsDelta.h"

/* CountNotZero (operators.h:9-16): a community is non-empty iff its total > 0. */
__host__ __device__ inline uint32_t count_not_zero(uint32_t tot) { return tot > 0 ? 1u : 0u; }

/* GetMaxValue (operators.h:30-40) applied left to right over candidates in
 * ascending community order keeps the lowest id among equal values; the hash
 * sweep's atomicMax on (bits(dq) << 32 | c) (phaseOptimizationHash.cu:29-31)
 * keeps the highest.  `highest` selects the latter.  A candidate (s, c) is
 * better than the incumbent (bs, bc). */
template <typename S>
__host__ __device__ inline bool better_candidate(S s, uint32_t c, S bs, uint32_t bc, bool highest) {
    return s > bs || (s == bs && (highest ? c > bc : c < bc));
}

/* Square (operators.h:62-70), exact in 64 bits for totals < 2^31. */
__host__ __device__ inline uint64_t square_u64(uint32_t x) { return (uint64_t)x * (uint64_t)x; }

#endif
