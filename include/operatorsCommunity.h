/*
 * operatorsCommunity.h -- per-entry predicates and id maps.
 * Reference: operatorsCommunity.h:7-79 (TestTupleValue, MakeCommunityDest,
 * MakeCommunityPair, MakeCommunityBest, ActualNeighboorhood): the inline
 * rules the kernels use, and the functor structs under the reference's names.
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

/* ---- the reference's functors under their own names (:7-79).  Members and
 * constructors as in the reference; a thrust::tuple argument becomes an
 * argument list, a tuple result one of these structs. */

/* (source, community of the destination, weight) */
struct CommunityDest {
    unsigned int source;
    unsigned int community;
    float weight;
};

/* (community of the source, community of the destination, weight), renumbered */
struct CommunityPair {
    unsigned int source;
    unsigned int destination;
    unsigned weight;
};

/* TestTupleValue (:7-19) */
struct TestTupleValue {
    bool *changed;
    __host__ __device__ bool operator()(unsigned int source, unsigned int dest) const {
        return selected_entry(changed[source], source, dest);
    }
    TestTupleValue(bool *c) : changed(c) {}
};

/* MakeCommunityDest (:21-34): values = the communities */
struct MakeCommunityDest {
    unsigned int *values;
    __host__ __device__ CommunityDest operator()(unsigned int source, unsigned int dest,
                                                 unsigned w) const {
        return CommunityDest{source, values[dest], (float)w};
    }
    MakeCommunityDest(unsigned int *v) : values(v) {}
};

/* MakeCommunityPair (:37-51): map = inclusive scan of non-empty flags */
struct MakeCommunityPair {
    unsigned int *communities;
    unsigned int *map;
    __host__ __device__ CommunityPair operator()(unsigned int source, unsigned int dest,
                                                 unsigned w) const {
        return CommunityPair{renumbered(map, communities[source]),
                             renumbered(map, communities[dest]), w};
    }
    MakeCommunityPair(unsigned int *c, unsigned int *v) : communities(c), map(v) {}
};

/* MakeCommunityBest (:54-64) */
struct MakeCommunityBest {
    unsigned int *map;
    unsigned int *communities;
    __host__ __device__ unsigned int operator()(unsigned int d) const {
        return renumbered(map, communities[d]);
    }
    MakeCommunityBest(unsigned int *c, unsigned int *v) : map(v), communities(c) {}
};

/* ActualNeighboorhood (:67-79): over MakeCommunityDest's (source, community
 * of the destination, weight) */
struct ActualNeighboorhood {
    unsigned int *communities;
    __host__ __device__ bool operator()(unsigned int source, unsigned int community_dest,
                                        unsigned) const {
        return internal_entry(communities[source], community_dest);
    }
    ActualNeighboorhood(unsigned int *c) : communities(c) {}
};

#endif
