/*
 * operatorsDelta.h -- delta-modularity of moving vertex v to community c.
 *
 * Reference: DeltaModularityHash / DeltaModularitySort (operatorsDelta.h:7-45):
 *   float value = ((w_vc - self) / m) + ((k_v * (tot_own - k_v - tot_c)) / (2*m*m));
 * with the first difference in f32, everything else in f64, rounded to f32.
 * Here the same quantity is available two ways:
 *   delta_modularity_ref_f32  -- the reference's rounding, step for step
 *                                (delta_modularity_ref with the reference's
 *                                f32 weights and f64 totals);
 *   delta_modularity_exact    -- 2m^2 * delta-Q as an exact int64:
 *                                2m*(w_vc - self) + k_v*(tot_own - k_v - tot_c).
 * Within one vertex the exact form orders candidates by
 *   move_score_exact = 2m*w_vc - k_v*tot_c
 * (the remaining terms are constant for v).  Requires 2m < 2^31.
 */
#ifndef LV_OPERATORS_DELTA_H
#define LV_OPERATORS_DELTA_H

#include <hip/hip_runtime.h>
#include <cstdint>

__host__ __device__ inline float delta_modularity_ref_f32(uint32_t w_vc, uint32_t self, uint32_t k,
                                                          uint32_t tot_own, uint32_t tot_c,
                                                          double total_weight) {
    const float d = (float)w_vc - (float)self;
    const double t1 = (double)d / total_weight;
    const double t2 = ((double)k * (((double)tot_own - (double)k) - (double)tot_c)) /
                      (2 * total_weight * total_weight);
    return (float)(t1 + t2);
}

/* The same expression with the reference's operand types: summed weights in
 * f32 (w_vc, self), totals in f64. */
__host__ __device__ inline float delta_modularity_ref(float w_vc, float self, unsigned int k,
                                                      double tot_own, double tot_c,
                                                      double total_weight) {
    const double gain = (double)(w_vc - self) / total_weight;
    const double loss = ((double)k * ((tot_own - (double)k) - tot_c)) /
                        (2 * total_weight * total_weight);
    return (float)(gain + loss);
}

/* DeltaModularityHash (operatorsDelta.h:7-26): the functor over one hash-table
 * slot, key = v << 32 | c with summed weight w.  Members and constructor as in
 * the reference; the thrust::tuple argument becomes (key, w). */
struct DeltaModularityHash {
    double *community_weight;
    unsigned *nodes_weight;
    double total_weight;
    float *self_community;
    unsigned int *communities;

    __host__ __device__ float operator()(unsigned long long key, float w) const {
        const unsigned int node = (unsigned int)(key >> 32);
        const unsigned int community = (unsigned int)key;
        return delta_modularity_ref(w, self_community[node], nodes_weight[node],
                                    community_weight[communities[node]],
                                    community_weight[community], total_weight);
    }

    DeltaModularityHash(double *w, unsigned *n, double m, float *s, unsigned int *c)
        : community_weight(w), nodes_weight(n), total_weight(m), self_community(s), communities(c) {}
};

/* DeltaModularitySort (operatorsDelta.h:29-45): the same over one reduced
 * (node, community, w) run; the tuple becomes (node, community, w). */
struct DeltaModularitySort {
    double *community_weight;
    unsigned *nodes_weight;
    double total_weight;
    float *self_community;
    unsigned int *communities;

    __host__ __device__ float operator()(unsigned int node, unsigned int community, float w) const {
        return delta_modularity_ref(w, self_community[node], nodes_weight[node],
                                    community_weight[communities[node]],
                                    community_weight[community], total_weight);
    }

    DeltaModularitySort(double *w, unsigned *n, double m, float *s, unsigned int *c)
        : community_weight(w), nodes_weight(n), total_weight(m), self_community(s), communities(c) {}
};

__host__ __device__ inline int64_t move_score_exact(uint32_t w_vc, uint32_t k, uint32_t tot_c,
                                                    int64_t m2) {
    return m2 * (int64_t)w_vc - (int64_t)k * (int64_t)tot_c;
}

/* 2m^2 * delta-Q given the candidate's move score. */
__host__ __device__ inline int64_t delta_modularity_exact(int64_t score, uint32_t self, uint32_t k,
                                                          uint32_t tot_own, int64_t m2) {
    return score - m2 * (int64_t)self + (int64_t)k * ((int64_t)tot_own - (int64_t)k);
}

#endif
