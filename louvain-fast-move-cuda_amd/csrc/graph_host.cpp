// graph_host.cpp -- GraphHost: .edge parsing and host CSR construction
// (reference: graph.cu:17-87 for parsing, graph.cuh:46-99 for the sort by
// (src,dst), the min_id shift and the degree computation).
#include "../../include/graphHost.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>

#include "../../include/device_array.h"
#include "host_sort.h"

namespace {

// `std::istream >> int` token rules: whitespace, optional sign, digits.
bool read_int(const char *&p, const char *end, int64_t &out) {
    const char *q = p;
    while (q < end && (*q == ' ' || *q == '\t' || *q == '\n' || *q == '\r' || *q == '\v' ||
                       *q == '\f'))
        ++q;
    bool neg = false;
    if (q < end && (*q == '-' || *q == '+')) { neg = *q == '-'; ++q; }
    if (q >= end || *q < '0' || *q > '9') return false;
    int64_t v = 0;
    while (q < end && *q >= '0' && *q <= '9') {
        v = v * 10 + (*q - '0');
        if (v > (int64_t)std::numeric_limits<int>::max() + 1) return false;
        ++q;
    }
    v = neg ? -v : v;
    if (v > std::numeric_limits<int>::max() || v < std::numeric_limits<int>::min()) return false;
    out = v;
    p = q;
    return true;
}

}  // namespace

GraphHost::GraphHost(std::string name, bool weighted, int skip_line) {
    std::ifstream f(name, std::ios::binary);
    if (!f) throw lv::Error(LV_ERR_IO, "cannot open " + name);
    std::string buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const char *p = buf.data(), *end = buf.data() + buf.size();
    for (int i = 0; i < skip_line && p < end; ++i) {           // graph.cu:22-25
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
    std::vector<int64_t> a, b;
    std::vector<uint32_t> w;
    for (;;) {                                                 // graph.cu:36-44
        int64_t x, y, z = 1;
        const char *q = p;
        if (!read_int(q, end, x) || !read_int(q, end, y)) break;
        if (weighted && (!read_int(q, end, z) || z < 0)) break;
        p = q;
        a.push_back(x);
        b.push_back(y);
        if (weighted) w.push_back((uint32_t)z);
    }
    *this = from_edges(a.size(), a.data(), b.data(), weighted ? w.data() : nullptr);
}

GraphHost GraphHost::from_edges(uint64_t m, const int64_t *a, const int64_t *b, const uint32_t *w) {
    if (m == 0) throw lv::Error(LV_ERR_PARSE, "no edge parsed");  // graph.cu:52-55
    if (m > 0xFFFFFFFFull) throw lv::Error(LV_ERR_RANGE, "more than 2^32-1 edges");
    int64_t lo = std::numeric_limits<int64_t>::max(), hi = std::numeric_limits<int64_t>::min();
    for (uint64_t i = 0; i < m; ++i) {
        const int64_t x = std::min(a[i], b[i]), y = std::max(a[i], b[i]);
        lo = std::min(lo, x);
        hi = std::max(hi, y);
    }
    if (hi - lo >= (int64_t)0xFFFFFFFFll) throw lv::Error(LV_ERR_RANGE, "id span >= 2^32-1");
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    // directed entries: non-loops both ways, loops once (expanded to two below);
    // value = input index so that the first occurrence of a duplicate wins
    uint64_t loops = 0;
    for (uint64_t i = 0; i < m; ++i) loops += a[i] == b[i];
    const uint64_t ne = 2 * m - loops;
    std::vector<uint64_t> key(ne);
    std::vector<uint32_t> idx(ne);
    {
        uint64_t o = 0;
        for (uint64_t i = 0; i < m; ++i) {
            const uint64_t x = (uint64_t)(std::min(a[i], b[i]) - lo);
            const uint64_t y = (uint64_t)(std::max(a[i], b[i]) - lo);
            key[o] = (x << 32) | y; idx[o] = (uint32_t)i; ++o;
            if (x != y) { key[o] = (y << 32) | x; idx[o] = (uint32_t)i; ++o; }
        }
    }
    lv::radix_sort_pairs_host(key, &idx);                     // graph.cuh:60-64 (stable)
    GraphHost g;
    std::vector<uint8_t> seen(span, 0);
    std::vector<uint64_t> deg(span, 0);
    uint64_t u = 0, n_loops = 0, n_pairs = 0;
    for (uint64_t j = 0; j < ne; ++j) {
        if (j > 0 && key[j] == key[j - 1]) continue;          // std::set dedupe, graph.cu:26-43
        key[u] = key[j];
        idx[u] = idx[j];
        const uint64_t s = key[j] >> 32, d = key[j] & 0xFFFFFFFFull;
        seen[s] = seen[d] = 1;
        if (s == d) { deg[s] += 2; ++n_loops; }                // (a,a) appears twice, graph.cu:68-79
        else { deg[s] += 1; ++n_pairs; }
        ++u;
    }
    for (uint64_t i = 0; i < span; ++i)
        if (!seen[i]) throw lv::Error(LV_ERR_NOT_DENSE, "vertex ids are not dense");
    g.n_nodes = (unsigned)span;
    g.min_id = lo;
    g.n_links = n_loops + n_pairs / 2;
    g.row.assign(span + 1, 0);
    for (uint64_t v = 0; v < span; ++v) g.row[v + 1] = g.row[v] + deg[v];
    const uint64_t E = g.row[span];
    g.col.resize(E);
    g.unit_weights = (w == nullptr);
    if (w) {
        g.unit_weights = true;
        for (uint64_t j = 0; j < u; ++j)
            if (w[idx[j]] != 1u) { g.unit_weights = false; break; }
    }
    if (!g.unit_weights) g.weights.resize(E);
    uint64_t o = 0;
    double wsum = 0;
    for (uint64_t j = 0; j < u; ++j) {
        const uint32_t s = (uint32_t)(key[j] >> 32), d = (uint32_t)(key[j] & 0xFFFFFFFFull);
        const uint32_t wj = w ? w[idx[j]] : 1u;
        const int reps = s == d ? 2 : 1;
        for (int r = 0; r < reps; ++r) {
            g.col[o] = d;
            if (!g.unit_weights) g.weights[o] = wj;
            ++o;
        }
        wsum += (double)wj * reps;
    }
    g.total_weight = wsum / 2;                                // graph.cu:81 (= n_links for unit weights)
    return g;
}

// The reference's COO order (graph.cu:59-76): the set of (min, max) pairs in
// lexicographic order, each as (a, b) then (b, a).  Rows are sorted by
// destination, so the entries (v, x) with v <= x of the CSR walked in order
// are that set; a self-loop has two (v, v) entries and is emitted once.
void GraphHost::build_coo() {
    edge_source.clear();
    edge_destination.clear();
    edge_weights.clear();
    edge_source.reserve(col.size());
    edge_destination.reserve(col.size());
    edge_weights.reserve(col.size());
    for (uint32_t v = 0; v < n_nodes; ++v) {
        bool second_self = false;
        for (uint64_t j = row[v]; j < row[v + 1]; ++j) {
            const uint32_t x = col[j];
            if (x < v) continue;
            if (x == v) {
                second_self = !second_self;
                if (second_self) continue;
            }
            const uint32_t w = unit_weights ? 1u : weights[j];
            const int64_t a = (int64_t)v + min_id, b = (int64_t)x + min_id;
            edge_source.push_back(a); edge_destination.push_back(b); edge_weights.push_back(w);
            edge_source.push_back(b); edge_destination.push_back(a); edge_weights.push_back(w);
        }
    }
}

GraphHost GraphHost::from_csr(uint32_t n, uint64_t E, const uint64_t *row, const uint32_t *col,
                              const uint32_t *w, double total_weight) {
    if (row[0] != 0 || row[n] != E) throw lv::Error(LV_ERR_ARG, "CSR offsets inconsistent");
    GraphHost g;
    g.n_nodes = n;
    g.row.assign(row, row + n + 1);
    g.col.assign(col, col + E);
    double wsum = 0;
    g.unit_weights = true;
    if (w)
        for (uint64_t j = 0; j < E; ++j)
            if (w[j] != 1u) { g.unit_weights = false; break; }
    if (!g.unit_weights) g.weights.assign(w, w + E);
    for (uint32_t v = 0; v < n; ++v) {
        if (row[v + 1] < row[v]) throw lv::Error(LV_ERR_ARG, "CSR offsets decrease");
        for (uint64_t j = row[v]; j < row[v + 1]; ++j) {
            if (col[j] >= n) throw lv::Error(LV_ERR_ARG, "column id out of range");
            wsum += w ? w[j] : 1.0;
        }
    }
    // rows in destination order, as the reference's sort by (src, dst) leaves them
    std::vector<std::pair<uint32_t, uint32_t>> tmp;
    for (uint32_t v = 0; v < n; ++v) {
        const uint64_t b = g.row[v], e = g.row[v + 1];
        bool sorted = true;
        for (uint64_t j = b + 1; j < e; ++j)
            if (g.col[j] < g.col[j - 1]) { sorted = false; break; }
        if (sorted) continue;
        tmp.clear();
        for (uint64_t j = b; j < e; ++j) tmp.emplace_back(g.col[j], g.unit_weights ? 1u : g.weights[j]);
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const auto &x, const auto &y) { return x.first < y.first; });
        for (uint64_t j = b; j < e; ++j) {
            g.col[j] = tmp[j - b].first;
            if (!g.unit_weights) g.weights[j] = tmp[j - b].second;
        }
    }
    // The exact gains use 2m = 2 * total_weight and rely on 2m being the
    // degree sum (whose bound 2m < 2^31 the community checks), so a caller's
    // total_weight must be half the weight sum, as it is for every graph the
    // reference builds (graph.cu:81, phaseAggregation.cuh:172).
    if (total_weight > 0 && 2.0 * total_weight != wsum)
        throw lv::Error(LV_ERR_ARG, "total_weight is not half the sum of the weights");
    g.total_weight = total_weight > 0 ? total_weight : wsum / 2;
    g.n_links = E / 2;
    g.min_id = 0;
    return g;
}
