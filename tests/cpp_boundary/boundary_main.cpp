// The reference's caller code compiled against include/ and linked to
// liblouvain_amd.so: main.cu:16-22 and :32-33 line for line (its
// `GraphHost::GraphHost(...)` spelling is an MSVC extension, written here in
// the standard form), the level loop of modularityAlgorithms.h:31-67 phase by
// phase, the functors of operators*.h under their reference names, and the
// COO members.  It prints key=value lines and writes the partitions next to
// the given prefix; tests/test_gpu_cpp_boundary.py compares them with the
// oracle.
//
//   boundary_main <graph.edge> <output prefix>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "modularityAlgorithms.h"
#include "operators.h"

static std::vector<uint32_t> to_host(const lv::DeviceArray<uint32_t> &a, size_t n) {
    std::vector<uint32_t> h(n);
    if (n && hipMemcpy(h.data(), a.data(), n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy");
    return h;
}

static void save(const std::string &path, const std::vector<uint32_t> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(uint32_t)));
}

static uint32_t f32_bits(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return b;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s graph.edge out_prefix\n", argv[0]);
        return 2;
    }
    const std::string name = argv[1], out = argv[2];
    try {
        // device-independent parts first (checked without a GPU too)
        {
            // the functors of operatorsDelta.h / operators.h / operatorsCommunity.h
            double community_weight[4] = {7, 3, 12, 5};
            unsigned nodes_weight[4] = {3, 2, 4, 1};
            float self_community[4] = {1, 0, 2, 0};
            unsigned communities[4] = {0, 1, 2, 2};
            const double m = 19996;
            DeltaModularityHash dh(community_weight, nodes_weight, m, self_community, communities);
            DeltaModularitySort ds(community_weight, nodes_weight, m, self_community, communities);
            // (node, community, w): node 0 (own 0) to 1 with w 2, node 2 (own 2) to 0 with w 5
            std::printf("dq_hash_0_1=%u\n", f32_bits(dh((0ull << 32) | 1u, 2.0f)));
            std::printf("dq_sort_2_0=%u\n", f32_bits(ds(2u, 0u, 5.0f)));
            GetMaxValue gm;
            CommunityValue vals[4] = {{4, 0.5f}, {9, 0.75f}, {2, 0.75f}, {7, 0.25f}};
            CommunityValue best = vals[0];
            for (int i = 1; i < 4; ++i) best = gm(best, vals[i]);   // left to right, as reduce_by_key
            std::printf("getmax_community=%u\n", best.community);
            unsigned map[4] = {1, 2, 2, 3};   // inclusive scan of (1, 1, 0, 1)
            bool changed[4] = {true, false, true, true};
            std::printf("count_not_zero=%d%d\n", CountNotZero()(2.0), CountNotZero()(0.0));
            std::printf("test_tuple=%d%d%d\n", TestTupleValue(changed)(0, 1), TestTupleValue(changed)(1, 2),
                        TestTupleValue(changed)(2, 2));
            const CommunityDest cd = MakeCommunityDest(communities)(1, 3, 6u);
            std::printf("make_dest=%u,%u,%g\n", cd.source, cd.community, (double)cd.weight);
            const CommunityPair cp = MakeCommunityPair(communities, map)(0, 3, 4u);
            std::printf("make_pair=%u,%u,%u\n", cp.source, cp.destination, cp.weight);
            std::printf("make_best=%u\n", MakeCommunityBest(communities, map)(3));
            std::printf("actual=%d%d\n", ActualNeighboorhood(communities)(3, 2, 1u),
                        ActualNeighboorhood(communities)(0, 2, 1u));
            const SplitEntry se = SplitIterator()((5ull << 32) | 9u, 1.5f);
            std::printf("split=%u,%u,%g\n", se.source, se.destination, (double)se.weight);
            const LvPair<unsigned, unsigned> ps = PairSum<unsigned, unsigned>()({1, 2}, {3, 4});
            std::printf("pair_sum=%u,%u\n", ps.first, ps.second);
            std::printf("square=%g\nequals=%d\n", Square<double>()(1.5), Equals<int>(3)(3) ? 1 : 0);
            std::fflush(stdout);
        }
        GraphHost g(name, false, 0);                                     // main.cu:16
        const uint32_t n0 = g.n_nodes;
        {
            // COO members: the host pairs in the reference's order
            g.build_coo();
            std::printf("coo_entries=%zu\n", g.edge_source.size());
            std::ofstream f(out + ".coo.txt");
            for (size_t i = 0; i < g.edge_source.size(); ++i)
                f << g.edge_source[i] << " " << g.edge_destination[i] << " " << g.edge_weights[i] << "\n";
            std::fflush(stdout);
        }
        {
            auto C = ModularityAlgorithms::Louvain(g, HASH);              // main.cu:21
            std::cout << "hash_n_comm=" << C.n_of_best_communities << std::endl;  // main.cu:22
            std::printf("hash_q=%.17g\n", C.modularity);
            save(out + ".hash.bin", to_host(C.best_communities, n0));
        }
        {
            auto C = ModularityAlgorithms::Louvain(g, SORT);              // main.cu:32
            std::cout << "sort_n_comm=" << C.n_of_best_communities << std::endl;  // main.cu:33
            std::printf("sort_q=%.17g\n", C.modularity);
            save(out + ".sort.bin", to_host(C.best_communities, n0));
        }
        {
            // modularity_routine (modularityAlgorithms.h:18-67), phase by phase
            Community C(g);
            double old_modularity = 0;
            bool continue_optimization = false;
            int levels = 0;
            do {
                old_modularity = C.modularity;
                OptimizationPhase::run(C, HASH);
                continue_optimization = (C.modularity - old_modularity) > MODULARITY_CONVERGED_THRESHOLD;
                AggregationPhase::run(C, continue_optimization, HASH);
                RefinePhase::run();
                ++levels;
            } while (continue_optimization);
            std::printf("loop_n_comm=%u\nloop_q=%.17g\nloop_levels=%d\n", C.n_of_best_communities,
                        C.modularity, levels);
            save(out + ".loop.bin", to_host(C.best_communities, n0));
        }
        {
            // COO members: the device sources
            GraphDevice d(g, nullptr);
            d.build_coo(nullptr);
            if (hipDeviceSynchronize() != hipSuccess) throw std::runtime_error("sync");
            const auto src = to_host(d.edge_source, d.n_entries);
            const auto deg = to_host(d.n_of_neighboor, d.n_nodes);
            bool ok = true;
            uint64_t j = 0;
            for (uint32_t v = 0; v < d.n_nodes; ++v) {
                ok = ok && deg[v] == g.row[v + 1] - g.row[v];
                for (uint32_t k = 0; k < deg[v]; ++k, ++j) ok = ok && j < src.size() && src[j] == v;
            }
            std::printf("device_coo_ok=%d\n", ok && j == d.n_entries ? 1 : 0);
        }
    } catch (const lv::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
