// Host AddressSanitizer/UBSan driver for the library's host code and the
// oracle: .edge parsing (graph_host.cpp) on the ingest edge cases and
// generated files, CSR construction and its errors, the generators
// (generators.cpp), and oracle runs (oracle/louvain_oracle.c) in every mode.
// Built with g++/gcc only (tests/asan/Makefile); no GPU code is involved.
//
//   asan_driver <bundled fixture .edge> <scratch dir>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/graphHost.h"
#include "../../include/lv_api.h"
#include "../../include/device_array.h"

extern "C" {
#include "../../oracle/louvain_oracle.h"
}

static int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

static int parse_status(const std::string &path, bool weighted, int skip) {
    try {
        GraphHost g(path, weighted, skip);
        g.build_coo();
        CHECK(g.edge_source.size() == g.edge_destination.size());
        return 0;
    } catch (const lv::Error &e) {
        return (int)e.status;
    }
}

static void write(const std::string &path, const std::string &text) {
    std::ofstream f(path, std::ios::binary);
    f << text;
}

static void oracle_runs(lvo_graph *g) {
    std::vector<uint32_t> best(g->n);
    std::vector<lvo_sweep_rec> sw(4096);
    std::vector<lvo_level_rec> lv(256);
    for (int mode = 0; mode < 4; ++mode)
        for (int gain = 0; gain < 2; ++gain)
            for (int tie = 0; tie < 2; ++tie)
                for (int chunk = 0; chunk < 2; ++chunk) {
                    lvo_params p;
                    lvo_default_params(&p);
                    p.mode = mode;
                    p.gain = gain;
                    p.tie = tie;
                    p.chunk_emulation = chunk;
                    p.chunk_entries = 5000;
                    p.refine = chunk;
                    lvo_result r = {};
                    r.best_communities = best.data();
                    r.sweeps = sw.data();
                    r.sweeps_cap = (int32_t)sw.size();
                    r.level_recs = lv.data();
                    r.levels_cap = (int32_t)lv.size();
                    CHECK(lvo_louvain(g, &p, &r) == LVO_OK);
                    CHECK(r.n_of_best_communities >= 1 && r.n_of_best_communities <= g->n);
                }
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string fixture = argv[1], dir = argv[2];
    // parsing edge cases (graph.cu:17-87 semantics)
    const char *cases[][2] = {
        {"1 2\n2 3\n3 1\n", "0"},           {"  1\t2\r\n2 3\r\n\r\n3 1", "0"},
        {"1 2\n2 3\nabc\n3 4\n", "0"},      {"1 2 3", "0"},
        {"1-2 3 4\n", "5"},                 {"+1 +2\n-1 1\n0 2\n", "0"},
        {"- 5\n1 2\n", "3"},                {"99999999999 1\n", "3"},
        {"1 2\n2147483648 1\n", "0"},       {"-2147483648 -2147483647\n", "0"},
        {"1 1\n1 2\n2 1\n1 1\n2 2\n", "0"}, {"5 6\n6 7", "0"},
        {"1 2\n5 6\n", "5"},                {"", "3"},
        {"hello\n", "3"},
    };
    int i = 0;
    for (auto &c : cases) {
        const std::string p = dir + "/case" + std::to_string(i++) + ".edge";
        write(p, c[0]);
        CHECK(parse_status(p, false, 0) == std::stoi(c[1]));
        lvo_graph *og = nullptr;
        const int orc = lvo_graph_load_edge(p.c_str(), 0, 0, &og);
        CHECK(orc == std::stoi(c[1]));
        lvo_graph_free(og);
    }
    write(dir + "/w.edge", "c x\n1 2 5\n2 3 7\n1 2 9\n3 1 -1\n2 2 3\n");
    CHECK(parse_status(dir + "/w.edge", true, 1) == 0);
    CHECK(parse_status(dir + "/missing.edge", false, 0) == 4);
    // the fixture through both readers, then every oracle mode on it
    CHECK(parse_status(fixture, false, 0) == 0);
    lvo_graph *fg = nullptr;
    CHECK(lvo_graph_load_edge(fixture.c_str(), 0, 0, &fg) == LVO_OK);
    if (fg) {
        oracle_runs(fg);
        std::vector<uint32_t> comm(fg->n), nid(fg->n), target(fg->n);
        std::vector<int64_t> tot(fg->n, 0);
        std::vector<uint8_t> act(fg->n, 1);
        for (uint32_t v = 0; v < fg->n; ++v) comm[v] = v % 97;
        for (uint32_t v = 0; v < fg->n; ++v) tot[comm[v]] += (int64_t)fg->k[v];
        lvo_params p;
        lvo_default_params(&p);
        for (int kind = 1; kind < 3; ++kind) {
            uint64_t ea = 0;
            uint32_t nc = 0;
            lvo_best_moves(fg, comm.data(), tot.data(), act.data(), kind, &p, target.data(), &ea, &nc);
        }
        uint32_t n_new = 0;
        lvo_graph *h = nullptr;
        CHECK(lvo_aggregate(fg, comm.data(), nid.data(), &n_new, &h) == LVO_OK && n_new == 97);
        lvo_graph_free(h);
        lvo_graph_free(fg);
    }
    // generators, the host CSR from their edges, and an oracle run on each
    for (int kind = 0; kind < 3; ++kind) {
        uint64_t m = 0;
        uint32_t n = 0;
        std::vector<int64_t> a, b;
        std::vector<uint32_t> w, planted;
        if (kind == 0) {
            CHECK(lv_gen_rmat(12, 8, 0.57, 0.19, 0.19, 7, nullptr, nullptr, 0, &m, &n) == LV_OK);
            a.resize(m); b.resize(m);
            CHECK(lv_gen_rmat(12, 8, 0.57, 0.19, 0.19, 7, a.data(), b.data(), m, &m, &n) == LV_OK);
        } else if (kind == 1) {
            CHECK(lv_gen_ba(3000, 3, 9, nullptr, nullptr, 0, &m) == LV_OK);
            a.resize(m); b.resize(m);
            CHECK(lv_gen_ba(3000, 3, 9, a.data(), b.data(), m, &m) == LV_OK);
        } else {
            planted.resize(4000);
            CHECK(lv_gen_planted(4000, 2.5, 12, 200, 20, 1.5, 50, 400, 0.3, 1, 10, 3, nullptr,
                                 nullptr, nullptr, 0, &m, planted.data()) == LV_OK);
            a.resize(m); b.resize(m); w.resize(m);
            CHECK(lv_gen_planted(4000, 2.5, 12, 200, 20, 1.5, 50, 400, 0.3, 1, 10, 3, a.data(),
                                 b.data(), w.data(), m, &m, planted.data()) == LV_OK);
        }
        try {
            GraphHost g = GraphHost::from_edges(m, a.data(), b.data(), w.empty() ? nullptr : w.data());
            GraphHost c = GraphHost::from_csr(g.n_nodes, g.col.size(), g.row.data(), g.col.data(),
                                              g.unit_weights ? nullptr : g.weights.data(), 0.0);
            CHECK(c.total_weight == g.total_weight);
        } catch (const lv::Error &e) {
            CHECK(false);
        }
        const std::string p = dir + "/gen" + std::to_string(kind) + ".edge";
        CHECK(lv_write_edge(p.c_str(), a.data(), b.data(), w.empty() ? nullptr : w.data(), m) == LV_OK);
        CHECK(parse_status(p, !w.empty(), 0) == 0);
        lvo_graph *og = nullptr;
        CHECK(lvo_graph_from_edges(m, a.data(), b.data(), w.empty() ? nullptr : w.data(), &og) == LVO_OK);
        if (og) {
            oracle_runs(og);
            lvo_graph_free(og);
        }
    }
    // keys with three populated 16-bit digits (ids past 65,535, every pair's
    // smaller end below it): the oracle's radix sort makes an odd number of
    // passes and copies its result back from the scratch arrays
    {
        const uint32_t n = 70001;
        std::vector<int64_t> a(n - 1), b(n - 1);
        a[0] = 0; b[0] = 1;
        for (uint32_t y = 2; y < n; ++y) { a[y - 1] = 1; b[y - 1] = y; }
        lvo_graph *og = nullptr;
        CHECK(lvo_graph_from_edges(n - 1, a.data(), b.data(), nullptr, &og) == LVO_OK);
        if (og) {
            CHECK(og->n == n && og->E == 2ull * (n - 1));
            CHECK(og->row[1] == 1 && og->row[2] == n);   // vertex 1 meets every other vertex
            lvo_graph_free(og);
        }
    }
    // CSR errors
    const uint64_t row_bad[3] = {0, 2, 1};
    const uint32_t col_bad[2] = {1, 5};
    try {
        GraphHost::from_csr(2, 1, row_bad, col_bad, nullptr, 0.0);
        CHECK(false);
    } catch (const lv::Error &e) {
        CHECK(e.status == LV_ERR_ARG);
    }
    std::printf("asan driver: %d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
