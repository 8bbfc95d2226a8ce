/*
 * graph.h -- one level's CSR resident in HBM (reference: GraphDevice,
 * graph.cuh:14-102).  Field names follow the reference where the meaning is
 * the same: neighboorhood_sum = row offsets (here with the terminal entry),
 * edge_destination = column ids, weights, tot_weight_per_nodes = k_v.
 */
#ifndef LV_GRAPH_H
#define LV_GRAPH_H

#include <hip/hip_runtime.h>

#include "device_array.h"
#include "graphHost.h"

struct GraphDevice {
    unsigned int n_nodes = 0;
    unsigned long n_links = 0;
    double total_weight = 0;       // m of the level-0 graph, never updated (phaseAggregation.cuh:172)
    uint64_t n_entries = 0;
    uint64_t m2 = 0;               // sum of k_v = 2m
    bool unit_weights = true;

    lv::DeviceArray<uint64_t> neighboorhood_sum;     // [n+1]
    lv::DeviceArray<uint32_t> edge_destination;      // [E]
    lv::DeviceArray<uint32_t> weights;               // [E], empty when unit
    lv::DeviceArray<uint32_t> tot_weight_per_nodes;  // [n]
    // The reference's COO members (graph.cuh:25-32): the source of every
    // entry (entries in CSR order, so edge_destination is the same array) and
    // every vertex's entry count.  The optimisation loop does not use them;
    // they stay empty until build_coo() fills them.
    lv::DeviceArray<uint32_t> edge_source;           // [E]
    lv::DeviceArray<uint32_t> n_of_neighboor;        // [n]

    GraphDevice() = default;
    GraphDevice(const GraphHost &g, hipStream_t s);  // graph.cuh:46-99
    GraphDevice(GraphDevice &&) = default;
    GraphDevice &operator=(GraphDevice &&) = default;
    // Non-owning view of another level graph (the level-0 CSR kept for re-runs).
    static GraphDevice view_of(const GraphDevice &g);
    // Fills edge_source and n_of_neighboor from the CSR (kernels on stream s).
    void build_coo(hipStream_t s);
};

#endif
