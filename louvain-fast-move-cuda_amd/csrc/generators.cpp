// generators.cpp -- seeded synthetic graphs for the measured workloads.
//
// Random numbers come from a counter-based mixer (SplitMix64 finaliser over
// seed, stream and counter), so every output depends only on the seed, never
// on the thread count.  The reference ships one bundled input
// (graph-power-law-10000-2-99-1.edge); these generators produce the R-MAT,
// Barabasi-Albert and planted-partition inputs of the measured workloads.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/device_array.h"
#include "../../include/lv_api.h"
#include "host_sort.h"

namespace {

inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed * 0x9E3779B97F4A7C15ull + stream) + ctr * 0xD1B54A32D192ED03ull);
}

inline double rng_unit(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return (double)(rng_u64(seed, stream, ctr) >> 11) * (1.0 / 9007199254740992.0);
}

// Sort + unique undirected keys (x<<32|y, x<y); compact the ids of vertices
// that keep an edge in increasing order; write the edge list.
lv_status finish_edges(std::vector<uint64_t> &keys, int64_t *ea, int64_t *eb, uint64_t cap,
                       uint64_t *m, uint32_t *n_out, uint64_t id_space) {
    lv::radix_sort_pairs_host(keys, nullptr);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<uint32_t> newid(id_space, 0);
    for (uint64_t k : keys) { newid[k >> 32] = 1; newid[k & 0xFFFFFFFFull] = 1; }
    uint32_t run = 0;
    for (uint64_t i = 0; i < id_space; ++i) {
        const uint32_t used = newid[i];
        newid[i] = run;
        run += used;
    }
    if (n_out) *n_out = run;
    *m = keys.size();
    if (!ea || !eb) return LV_OK;
    if (cap < keys.size()) return LV_ERR_ARG;
    lv::parallel_for_chunks(keys.size(), lv::host_threads(), [&](int, size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            ea[i] = newid[keys[i] >> 32];
            eb[i] = newid[keys[i] & 0xFFFFFFFFull];
        }
    });
    return LV_OK;
}

}  // namespace

extern "C" {

lv_status lv_gen_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                      int64_t *ea, int64_t *eb, uint64_t cap, uint64_t *m, uint32_t *n_out) {
    if (!m || scale < 1 || scale > 31 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0)
        return LV_ERR_ARG;
    try {
        const uint64_t nv = 1ull << scale;
        const uint64_t draws = (uint64_t)edge_factor << scale;
        const double ab = a + b, abc = a + b + c;
        std::vector<uint64_t> keys(draws);
        std::vector<uint8_t> keep(draws);
        lv::parallel_for_chunks(draws, lv::host_threads(), [&](int, size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) {
                uint64_t u = 0, v = 0;
                for (int l = 0; l < scale; ++l) {
                    const double r = rng_unit(seed, i, (uint64_t)l);
                    const uint64_t bit = 1ull << (scale - 1 - l);
                    if (r < a) {
                    } else if (r < ab) {
                        v |= bit;
                    } else if (r < abc) {
                        u |= bit;
                    } else {
                        u |= bit;
                        v |= bit;
                    }
                }
                keep[i] = u != v;   // drop self-loops
                const uint64_t x = std::min(u, v), y = std::max(u, v);
                keys[i] = (x << 32) | y;
            }
        });
        size_t o = 0;
        for (size_t i = 0; i < draws; ++i)
            if (keep[i]) keys[o++] = keys[i];
        keys.resize(o);
        return finish_edges(keys, ea, eb, cap, m, n_out, nv);
    } catch (const std::bad_alloc &) {
        return LV_ERR_OOM;
    }
}

// Preferential attachment: vertex s >= m_attach joins with m_attach distinct
// targets drawn uniformly from the list of edge endpoints so far (the first
// joiner takes the m_attach seed vertices).
lv_status lv_gen_ba(uint32_t n, uint32_t m_attach, uint64_t seed, int64_t *ea, int64_t *eb,
                    uint64_t cap, uint64_t *m) {
    if (!m || m_attach < 1 || n <= m_attach) return LV_ERR_ARG;
    const uint64_t total = (uint64_t)(n - m_attach) * m_attach;
    *m = total;
    if (!ea || !eb) return LV_OK;
    if (cap < total) return LV_ERR_ARG;
    try {
        std::vector<uint32_t> endpoints;
        endpoints.reserve(2 * total);
        std::vector<uint32_t> targets(m_attach);
        std::iota(targets.begin(), targets.end(), 0u);
        std::vector<uint32_t> mark(n, 0);
        uint64_t o = 0;
        for (uint32_t s = m_attach; s < n; ++s) {
            for (uint32_t t : targets) {
                ea[o] = s;
                eb[o] = t;
                ++o;
                endpoints.push_back(t);
                endpoints.push_back(s);
            }
            // choose m_attach distinct targets for the next vertex
            uint64_t ctr = 0;
            uint32_t got = 0;
            while (got < m_attach) {
                const uint64_t r = rng_u64(seed, s, ctr++);
                const uint32_t t = endpoints[r % endpoints.size()];
                if (mark[t] == s + 1) continue;
                mark[t] = s + 1;
                targets[got++] = t;
            }
        }
        return LV_OK;
    } catch (const std::bad_alloc &) {
        return LV_ERR_OOM;
    }
}

// Planted partition: community sizes ~ power law (size_exp) in [size_lo,
// size_hi] rescaled to sum to n; degrees ~ power law (deg_exp) with the lower
// cut solved for mean_deg and capped at max_deg; a fraction mu of each
// vertex's stubs is wired across communities, the rest inside
// (configuration model; self-loops and repeats dropped); weights uniform in
// [w_lo, w_hi].  Vertices left without an edge are removed and ids compacted
// (planted[] follows the compaction; *m counts edges).
lv_status lv_gen_planted(uint32_t n, double deg_exp, double mean_deg, uint32_t max_deg,
                         uint32_t n_comm, double size_exp, uint32_t size_lo, uint32_t size_hi,
                         double mu, uint32_t w_lo, uint32_t w_hi, uint64_t seed, int64_t *ea,
                         int64_t *eb, uint32_t *ew, uint64_t cap, uint64_t *m, uint32_t *planted) {
    if (!m || n < 2 || n_comm < 1 || size_lo < 1 || size_hi < size_lo || deg_exp <= 2.0 ||
        mean_deg <= 1.0 || max_deg < 2 || mu < 0 || mu > 1 || w_lo < 1 || w_hi < w_lo)
        return LV_ERR_ARG;
    try {
        // community sizes
        std::vector<double> sz(n_comm);
        const double lo = size_lo, hi = size_hi, ea1 = 1.0 - size_exp;
        for (uint32_t i = 0; i < n_comm; ++i) {
            const double u = rng_unit(seed, 1, i);
            sz[i] = std::pow(std::pow(lo, ea1) + u * (std::pow(hi, ea1) - std::pow(lo, ea1)), 1.0 / ea1);
        }
        const double ssum = std::accumulate(sz.begin(), sz.end(), 0.0);
        std::vector<uint32_t> size(n_comm);
        uint64_t acc = 0;
        for (uint32_t i = 0; i < n_comm; ++i) {
            size[i] = std::max<uint32_t>(1, (uint32_t)std::llround(sz[i] * n / ssum));
            acc += size[i];
        }
        // fix the total to n on the largest community
        const uint32_t big = (uint32_t)(std::max_element(size.begin(), size.end()) - size.begin());
        if (acc > n) size[big] -= (uint32_t)std::min<uint64_t>(acc - n, size[big] - 1);
        else size[big] += (uint32_t)(n - acc);
        std::vector<uint32_t> comm(n);
        {
            uint64_t v = 0;
            for (uint32_t i = 0; i < n_comm && v < n; ++i)
                for (uint32_t k = 0; k < size[i] && v < n; ++k) comm[v++] = i;
            while (v < n) comm[v++] = n_comm - 1;
        }
        // degree lower cut: mean of a power law truncated to [x0, max_deg]
        auto mean_of = [&](double x0) {
            const double g = deg_exp, X = max_deg;
            const double num = (std::pow(X, 2 - g) - std::pow(x0, 2 - g)) / (2 - g);
            const double den = (std::pow(X, 1 - g) - std::pow(x0, 1 - g)) / (1 - g);
            return num / den;
        };
        double x_lo = 1.0, x_hi = mean_deg;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (x_lo + x_hi);
            if (mean_of(mid) < mean_deg) x_lo = mid; else x_hi = mid;
        }
        const double x0 = 0.5 * (x_lo + x_hi);
        std::vector<uint32_t> din(n), dout(n);
        const double g1 = 1.0 - deg_exp;
        const double A = std::pow(x0, g1), B = std::pow((double)max_deg, g1);
        for (uint32_t v = 0; v < n; ++v) {
            const double u = rng_unit(seed, 2, v);
            double d = std::pow(A + u * (B - A), 1.0 / g1);
            uint32_t di = (uint32_t)std::llround(d);
            di = std::max<uint32_t>(1, std::min<uint32_t>(di, max_deg));
            const uint32_t ext = (uint32_t)std::llround(mu * di);
            dout[v] = ext;
            din[v] = di - ext;
        }
        std::vector<uint64_t> keys;
        auto pair_stubs = [&](std::vector<uint32_t> &stubs, uint64_t stream) {
            // Fisher-Yates with the counter-based generator, then pair neighbours
            for (size_t i = stubs.size(); i > 1; --i) {
                const size_t j = rng_u64(seed, stream, i) % i;
                std::swap(stubs[i - 1], stubs[j]);
            }
            for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
                const uint32_t x = stubs[i], y = stubs[i + 1];
                if (x == y) continue;
                keys.push_back(((uint64_t)std::min(x, y) << 32) | std::max(x, y));
            }
        };
        {
            uint64_t v0 = 0;
            std::vector<uint32_t> stubs;
            for (uint32_t ci = 0; ci < n_comm; ++ci) {
                stubs.clear();
                for (uint64_t v = v0; v < v0 + size[ci] && v < n; ++v)
                    for (uint32_t k = 0; k < din[v]; ++k) stubs.push_back((uint32_t)v);
                pair_stubs(stubs, 1000 + ci);
                v0 += size[ci];
            }
            stubs.clear();
            for (uint32_t v = 0; v < n; ++v)
                for (uint32_t k = 0; k < dout[v]; ++k) stubs.push_back(v);
            pair_stubs(stubs, 999);
        }
        uint32_t n_keep = 0;
        std::vector<uint64_t> kcopy = keys;
        uint64_t mm = 0;
        lv_status st = finish_edges(kcopy, nullptr, nullptr, 0, &mm, &n_keep, n);
        if (st != LV_OK) return st;
        *m = mm;
        if (!ea || !eb) return LV_OK;
        if (cap < mm) return LV_ERR_ARG;
        st = finish_edges(keys, ea, eb, cap, &mm, &n_keep, n);
        if (st != LV_OK) return st;
        // compaction map for planted labels (vertices keeping an edge, in order)
        std::vector<uint8_t> used(n, 0);
        for (uint64_t i = 0; i < keys.size(); ++i) {
            used[keys[i] >> 32] = 1;
            used[keys[i] & 0xFFFFFFFFull] = 1;
        }
        if (planted) {
            uint32_t o = 0;
            for (uint32_t v = 0; v < n; ++v)
                if (used[v]) planted[o++] = comm[v];
        }
        if (ew)
            for (uint64_t i = 0; i < mm; ++i)
                ew[i] = w_lo + (uint32_t)(rng_u64(seed, 3, i) % (uint64_t)(w_hi - w_lo + 1));
        return LV_OK;
    } catch (const std::bad_alloc &) {
        return LV_ERR_OOM;
    }
}

lv_status lv_write_edge(const char *path, const int64_t *a, const int64_t *b, const uint32_t *w,
                        uint64_t m) {
    if (!path || (m && (!a || !b))) return LV_ERR_ARG;
    auto put = [](char *p, int64_t v) {   // decimal digits of v at p; returns the end
        char tmp[24];
        int k = 0;
        const bool neg = v < 0;
        uint64_t u = neg ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
        do { tmp[k++] = (char)('0' + u % 10); u /= 10; } while (u);
        if (neg) *p++ = '-';
        while (k) *p++ = tmp[--k];
        return p;
    };
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint64_t per = (m + nt - 1) / nt;
    std::vector<std::string> buf(nt);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            const uint64_t i0 = std::min<uint64_t>(m, t * per), i1 = std::min<uint64_t>(m, i0 + per);
            std::string &out = buf[t];
            out.resize((i1 - i0) * 36);
            char *p = &out[0];
            for (uint64_t i = i0; i < i1; ++i) {
                p = put(p, a[i]);
                *p++ = ' ';
                p = put(p, b[i]);
                if (w) {
                    *p++ = ' ';
                    p = put(p, w[i]);
                }
                *p++ = '\n';
            }
            out.resize(p - &out[0]);
        });
    for (auto &x : th) x.join();
    FILE *f = std::fopen(path, "wb");
    if (!f) return LV_ERR_IO;
    for (auto &x : buf)
        if (std::fwrite(x.data(), 1, x.size(), f) != x.size()) {
            std::fclose(f);
            return LV_ERR_IO;
        }
    return std::fclose(f) == 0 ? LV_OK : LV_ERR_IO;
}

}  // extern "C"
