/*
 * graphHost.h -- host graph read from an .edge file (reference:
 * graphHost.cuh:7-18, graph.cu:17-87).
 *
 * Semantics kept from the reference: skip `skip_line` lines, read integer
 * pairs until the first parse failure, normalise every pair to (min,max),
 * drop duplicates, emit both directions (a self-loop yields two (a,a)
 * entries), sort by (src,dst), subtract min_id, total_weight = edge count.
 * The result is stored as a CSR (the reference keeps a COO and sorts it on the
 * device, graph.cuh:60-98).  Extensions: an optional third integer weight
 * column (first occurrence of a duplicate wins) and an error (LV_ERR_NOT_DENSE)
 * instead of out-of-bounds indexing when ids are not dense.
 */
#ifndef LV_GRAPH_HOST_H
#define LV_GRAPH_HOST_H

#include <cstdint>
#include <string>
#include <vector>

struct GraphHost {
    unsigned int n_nodes = 0;
    unsigned long n_links = 0;     // unique undirected edges (self-loops included)
    double total_weight = 0;       // m: sum of undirected edge weights
    long long min_id = 0;

    std::vector<uint64_t> row;     // n_nodes + 1 offsets
    std::vector<uint32_t> col;     // 2*n_links - (#self-loops) ... entries
    std::vector<uint32_t> weights; // empty when every weight is 1
    bool unit_weights = true;
    // several independent graphs side by side (a batch): 2m < 2^31 holds per
    // graph and is checked by the batch, not over the union
    bool disjoint_union = false;

    // The reference's host COO (graphHost.cuh:12-14): every unique undirected
    // pair (a, b), a <= b, in lexicographic order as (a, b) then (b, a), ids
    // as read (min_id not subtracted), and the weights in the same order.
    // Empty until build_coo() fills them; the library itself uses the CSR.
    std::vector<int64_t> edge_source;
    std::vector<int64_t> edge_destination;
    std::vector<uint32_t> edge_weights;

    GraphHost() = default;
    // graph.cu:17 -- throws lv::Error on failure
    GraphHost(std::string name, bool weighted, int skip_line);
    // The same construction from in-memory pairs (w may be null).
    static GraphHost from_edges(uint64_t m, const int64_t *a, const int64_t *b, const uint32_t *w);
    // Direct CSR; total_weight <= 0 means sum(w)/2.
    static GraphHost from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                              const uint32_t *w, double total_weight);
    uint64_t n_entries() const { return col.size(); }
    void build_coo();
};

#endif
