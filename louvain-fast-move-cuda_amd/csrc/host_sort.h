// host_sort.h -- multithreaded stable LSD radix sort of (u64 key, u32 value)
// pairs on the host, used to build CSRs from edge lists and to deduplicate
// generated edges.  16-bit digits; passes whose digit is zero in every key are
// skipped.  Stable: equal keys keep their input order.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace lv {

inline int host_threads() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    if (h > 16) h = 16;
    return (int)h;
}

template <typename F>
inline void parallel_for_chunks(size_t n, int threads, F &&f) {
    if (threads <= 1 || n < 65536) {
        f(0, (size_t)0, n);
        return;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) {
        const size_t b = n * t / threads, e = n * (t + 1) / threads;
        ts.emplace_back([&f, t, b, e] { f(t, b, e); });
    }
    for (auto &th : ts) th.join();
}

inline void radix_sort_pairs_host(std::vector<uint64_t> &key, std::vector<uint32_t> *val) {
    const size_t n = key.size();
    if (n < 2) return;
    const int T = n < 65536 ? 1 : host_threads();
    std::vector<uint64_t> orbits(T, 0);
    parallel_for_chunks(n, T, [&](int t, size_t b, size_t e) {
        uint64_t o = 0;
        for (size_t i = b; i < e; ++i) o |= key[i];
        orbits[t] = o;
    });
    uint64_t all = 0;
    for (auto o : orbits) all |= o;
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> v2(val ? n : 0);
    std::vector<uint64_t> cnt((size_t)T * 65536);
    uint64_t *ks = key.data(), *kd = k2.data();
    uint32_t *vs = val ? val->data() : nullptr, *vd = val ? v2.data() : nullptr;
    bool in_orig = true;
    for (int shift = 0; shift < 64; shift += 16) {
        if (((all >> shift) & 0xFFFFull) == 0) continue;
        std::fill(cnt.begin(), cnt.end(), 0);
        parallel_for_chunks(n, T, [&](int t, size_t b, size_t e) {
            uint64_t *c = cnt.data() + (size_t)t * 65536;
            for (size_t i = b; i < e; ++i) c[(ks[i] >> shift) & 0xFFFF]++;
        });
        uint64_t s = 0;
        for (int d = 0; d < 65536; ++d)
            for (int t = 0; t < T; ++t) {
                uint64_t &c = cnt[(size_t)t * 65536 + d];
                const uint64_t x = c;
                c = s;
                s += x;
            }
        parallel_for_chunks(n, T, [&](int t, size_t b, size_t e) {
            uint64_t *c = cnt.data() + (size_t)t * 65536;
            for (size_t i = b; i < e; ++i) {
                const uint64_t pos = c[(ks[i] >> shift) & 0xFFFF]++;
                kd[pos] = ks[i];
                if (vs) vd[pos] = vs[i];
            }
        });
        std::swap(ks, kd);
        std::swap(vs, vd);
        in_orig = !in_orig;
    }
    if (!in_orig) {
        std::memcpy(key.data(), ks, n * sizeof(uint64_t));
        if (val) std::memcpy(val->data(), vs, n * sizeof(uint32_t));
    }
}

}  // namespace lv
