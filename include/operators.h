/*
 * operators.h -- selection rules shared by the kernels.
 * Reference: operators.h:9-70 (CountNotZero, Equals, GetMaxValue, PairSum,
 * SplitIterator, Square): the inline rules the kernels use, and the
 * reference's functor structs under their own names.
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
    // both comparisons evaluated (bitwise, no short circuit): on the GPU a
    // short-circuit form becomes branches and execution-mask bookkeeping
    const bool tie = highest ? c > bc : c < bc;
    return (s > bs) | ((s == bs) & tie);
}

/* Square (operators.h:62-70), exact in 64 bits for totals < 2^31. */
__host__ __device__ inline uint64_t square_u64(uint32_t x) { return (uint64_t)x * (uint64_t)x; }

/* ---- the reference's functors under their own names (operators.h:9-70).
 * Members and constructors as in the reference; without Thrust, a
 * thrust::tuple argument becomes an argument list and a tuple result one of
 * the small structs below.  (MyUtils, the reference's debug printing, is not
 * part of this interface.) */

/* (community, value) pairs of the best-move reduction */
struct CommunityValue {
    unsigned int community;
    float value;
};

template <typename T1, typename T2>
struct LvPair {
    T1 first;
    T2 second;
};

/* (source, destination, weight) of one split hash key */
struct SplitEntry {
    unsigned int source;
    unsigned int destination;
    float weight;
};

/* CountNotZero (operators.h:9-16) */
struct CountNotZero {
    __host__ __device__ int operator()(double x) const { return x > 0 ? 1 : 0; }
    CountNotZero() {}
};

/* Equals (operators.h:18-28) */
template <typename T>
struct Equals {
    T value;
    __host__ __device__ bool operator()(T x) const { return x == value; }
    Equals(T v) : value(v) {}
};

/* GetMaxValue (operators.h:30-40): the later operand wins only if strictly
 * greater, so a left-to-right reduction keeps the first of equal values. */
struct GetMaxValue {
    __host__ __device__ CommunityValue operator()(CommunityValue a, CommunityValue b) const {
        return b.value > a.value ? b : a;
    }
    GetMaxValue() {}
};

/* PairSum (operators.h:42-49) */
template <typename T1, typename T2>
struct PairSum {
    __host__ __device__ LvPair<T1, T2> operator()(LvPair<T1, T2> p1, LvPair<T1, T2> p2) const {
        return LvPair<T1, T2>{p1.first + p2.first, p1.second + p2.second};
    }
};

/* SplitIterator (operators.h:51-60): key = a << 32 | b */
struct SplitIterator {
    __host__ __device__ SplitEntry operator()(unsigned long long key, float w) const {
        return SplitEntry{(unsigned int)(key >> 32), (unsigned int)key, w};
    }
    SplitIterator() {}
};

/* Square (operators.h:62-70) */
template <typename T>
struct Square {
    __host__ __device__ double operator()(T d) const { return d * d; }
    Square() {}
};

#endif
