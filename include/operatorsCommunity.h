/*
 * operatorsCommunity.h -- per-entry predicates and id maps.
 * Reference: operatorsCommunity.h:7-79 (TestTupleValue, MakeCommunityDest,
 * MakeCommunityPair, MakeCommunityBest, ActualNeighboorhood).
 */
#ifndef LV_OPERATORS_COMMUNITY_H
#define LV_OPERATORS_COMMUNITY_H

#include <hip/hip_runtime.h>
#include <cstdint>

/* TestTupleValue (:7-19): an entry takes part in a sweep iff its source is
 * active and it is not a self-loop. */
__host__ __device__ inline bool selected_entry(bool src_active, uint32_t src, uint32_t dst) {
    return src_active && src != dst;
}

/* MakeCommunityPair / MakeCommunityBest (:37-64): renumbered id of a
 * community from the inclusive-scan map (1-based). */
__host__ __device__ inline uint32_t renumbered(const uint32_t *map, uint32_t community) {
    return map[community] - 1u;
}

/* ActualNeighboorhood (:67-79): the entry lies inside one community. */
__host__ __device__ inline bool internal_entry(uint32_t comm_src, uint32_t comm_dst) {
    return comm_src == comm_dst;
}

#endif
