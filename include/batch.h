/*
 * batch.h -- many independent graphs through one launch sequence (no
 * counterpart in the reference, which runs one graph per process: main.cu:16-22).
 *
 * The graphs are laid side by side as one disjoint union (graph g's vertices
 * are the id range [voff[g], voff[g+1])).  Every kernel of a sweep or of an
 * aggregation then covers all graphs at once, while each graph keeps the
 * reference's own loop (modularityAlgorithms.h:14-81): its own 2m and m in the
 * gain, its own sweep-loop and level-loop convergence tests and counters, its
 * own modularity.  A graph whose sweep loop has ended is masked out of the
 * best-move pass until the level ends; a graph whose level loop has ended is
 * masked out for good and is carried through the remaining contractions
 * unchanged (each of its communities becomes one vertex that never moves), so
 * renumbering keeps its ids.  Per graph, the result equals that of
 * ModularityAlgorithms::Louvain on the graph alone.
 */
#ifndef LV_BATCH_H
#define LV_BATCH_H

#include <memory>
#include <vector>

#include "community.h"

struct GraphBatch {
    // graphs are copied into the union; 2m < 2^31 per graph
    explicit GraphBatch(const std::vector<const GraphHost *> &graphs, hipStream_t stream = nullptr);
    ~GraphBatch();
    GraphBatch(const GraphBatch &) = delete;
    GraphBatch &operator=(const GraphBatch &) = delete;

    // every graph from the identity partition to convergence
    void run(const lv_params &p, lv_batch_result *r);

    uint32_t n_graphs() const;
    uint64_t n_vertices() const;
    uint64_t n_entries() const;
    Community &union_community() { return *C_; }

  private:
    std::unique_ptr<Community> C_;
};

#endif
