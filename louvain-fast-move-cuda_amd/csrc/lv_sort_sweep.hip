// lv_sort_sweep.hip -- the sort variant of the best-move pass (reference:
// optimize_sort, phaseOptimizationSort.cu:161-379: select the entries of
// active vertices, sort them by (src, community of dst), reduce_by_key into
// w(v,c), take self = w(v, comm[v]), compute delta-Q per pair and keep the
// leftmost maximum per vertex).
//
// Here: every active vertex writes its row as keys (v << b | comm[x]) (self
// loops as a sentinel that sorts last) at a scanned offset; one LSD radix
// sort orders all keys; runs of equal keys are reduced with two scans; each
// active vertex then scans its own runs (sorted by community, so its own
// community is found by binary search), scores them and keeps the best with
// the lowest-id tie rule.  The decisions equal the hash variant's bit for bit.
#include "engine_state.h"
#include "lv_device.h"

namespace lv {

namespace {

struct SortCtx {
    BatchView bt;
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *wt;
    const uint32_t *comm;
    const uint32_t *tot;
    const uint32_t *kdeg;
    uint32_t *target;
    const uint64_t *off;      // [n] offset of each list position's row in the key array
    uint64_t *keys;
    uint32_t *vals;
    uint32_t *rbeg, *rend;    // [n] run range of each active vertex
    int bits;                 // community bits
    uint64_t sentinel;        // sorts after every real key
    uint32_t v_end;
    int64_t m2;
    double tw;
    bool highest;
    unsigned long long *stats;
};

template <int G, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_sort_gather(SortCtx a, const uint32_t *__restrict__ list,
                                                        uint32_t count, uint32_t list_base,
                                                        uint32_t parts) {
    constexpr int GPB = kBlock / G;
    const int lane = threadIdx.x & (G - 1);
    uint32_t idx, part;
    if constexpr (G == kBlock) {
        idx = blockIdx.x / parts;
        part = blockIdx.x % parts;
    } else {
        idx = blockIdx.x * GPB + group_index<G>();
        part = 0;
    }
    uint64_t nself = 0;
    if (idx < count) {
        const uint32_t v = list[idx];
        const uint64_t beg = a.row[v], end = a.row[v + 1];
        const uint64_t out = a.off[list_base + idx] - beg;
        if (lane == 0 && part == 0) { a.rbeg[v] = 0u; a.rend[v] = 0u; }
        int64_t m2;
        double tw;
        const bool live = v < a.v_end && vertex_graph(a.bt, v, a.m2, a.tw, m2, tw);
        const uint64_t stride = (uint64_t)G * (G == kBlock ? parts : 1u);
        for (uint64_t j = beg + (uint64_t)part * G + lane; j < end; j += stride) {
            const uint32_t x = stream_load(&a.col[j]);
            uint64_t key = a.sentinel;
            if (live && x != v) {
                key = ((uint64_t)v << a.bits) | a.comm[x];
                ++nself;
            }
            a.keys[out + j] = key;
            a.vals[out + j] = WEIGHTED ? stream_load(&a.wt[j]) : 1u;
        }
    }
    block_add_stats(a.stats, ST_E_ACTIVE, nself, -1, 0, -1, 0);
}

// Row degree of each list position (0 for positions past a bucket's count).
__global__ void k_sort_degrees(const uint64_t *row, const uint32_t *lists, uint32_t n,
                               const uint32_t *region_end, uint64_t *deg) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
         p += (size_t)gridDim.x * blockDim.x) {
        uint64_t d = 0;
        if (region_end[p] != 0u) {
            const uint32_t v = lists[p];
            d = row[v + 1] - row[v];
        }
        deg[p] = d;
    }
}

// Mark which list positions hold an active vertex (1) or a gap (0).
__global__ void k_sort_positions(uint32_t *valid, uint32_t n, uint64_t off0, uint32_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        valid[off0 + i] = 1u;
    (void)n;
}

__global__ void k_sort_ranges(const uint64_t *ukey, uint64_t nu, int bits, uint32_t n,
                              uint32_t *rbeg, uint32_t *rend) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nu;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = ukey[r] >> bits;
        if (v >= n) continue;   // sentinel run
        if (r == 0 || (ukey[r - 1] >> bits) != v) rbeg[v] = (uint32_t)r;
        if (r + 1 == nu || (ukey[r + 1] >> bits) != v) rend[v] = (uint32_t)(r + 1);
    }
}

// Group per active vertex over its runs: self by binary search, best candidate.
template <int G, int GAIN>
__global__ __launch_bounds__(kBlock) void k_sort_decide(SortCtx a, const uint32_t *__restrict__ list,
                                                        uint32_t count, const uint64_t *ukey,
                                                        const uint64_t *ws, const uint64_t *we) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    constexpr int GPB = kBlock / G;
    const int lane = threadIdx.x & (G - 1);
    const uint32_t idx = blockIdx.x * GPB + group_index<G>();
    const bool valid = idx < count;
    const uint32_t v = valid ? list[idx] : 0u;
    int64_t m2 = a.m2;
    double tw = a.tw;
    const bool live = valid && v < a.v_end && vertex_graph(a.bt, v, a.m2, a.tw, m2, tw);
    const uint64_t cmask = (1ull << a.bits) - 1;
    S best = SC::none();
    uint32_t best_c = a.highest ? 0u : LV_EMPTY;
    uint64_t ncand = 0;
    uint32_t self = 0, kv = 0, tot_own = 0, own = 0;
    uint32_t r0 = 0, r1 = 0;
    if (live) {
        r0 = a.rbeg[v];
        r1 = a.rend[v];
        own = a.comm[v];
        kv = a.kdeg[v];
        tot_own = a.tot[own];
        // select_self_node: the run (v, own), if any
        uint32_t lo = r0, hi = r1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint32_t)(ukey[mid] & cmask) < own) lo = mid + 1; else hi = mid;
        }
        if (lo < r1 && (uint32_t)(ukey[lo] & cmask) == own) self = (uint32_t)(we[lo] - ws[lo]);
        for (uint32_t r = r0 + lane; r < r1; r += G) {
            const uint32_t c = (uint32_t)(ukey[r] & cmask);
            if (c == own) continue;
            ++ncand;
            const uint32_t W = (uint32_t)(we[r] - ws[r]);
            const S sc = SC::of(W, self, kv, tot_own, a.tot[c], m2, tw);
            if (better_candidate(sc, c, best, best_c, a.highest)) { best = sc; best_c = c; }
        }
    }
    uint64_t considered = 0, twrite = 0;
    if constexpr (G <= 64) {
        group_best<G, S>(best, best_c, a.highest);
        if (live && lane == 0) {
            considered = r1 > r0;
            if (best != SC::none() && SC::positive(best, self, kv, tot_own, m2)) {
                a.target[v] = best_c;
                twrite = 1;
            }
        }
    } else {
        __shared__ S r_s[kBlock / 64];
        __shared__ uint32_t r_c[kBlock / 64];
        group_best<64, S>(best, best_c, a.highest);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; }
        __syncthreads();
        if (threadIdx.x == 0 && live) {
            for (int i = 0; i < kBlock / 64; ++i)
                if (better_candidate(r_s[i], r_c[i], best, best_c, a.highest)) {
                    best = r_s[i];
                    best_c = r_c[i];
                }
            considered = r1 > r0;
            if (best != SC::none() && SC::positive(best, self, kv, tot_own, m2)) {
                a.target[v] = best_c;
                twrite = 1;
            }
        }
    }
    block_add_stats(a.stats, ST_CANDIDATES, ncand, ST_CONSIDERED, considered, ST_TWRITES, twrite);
}

}  // namespace

static int bits_for_count(uint64_t x) {
    int b = 0;
    while (b < 64 && (1ull << b) < x) ++b;
    return b;
}

void sort_sweep(const SweepArgs &x, uint32_t *const *lists, const uint32_t *counts,
                const uint64_t *bucket_off, Workspace &ws, hipStream_t s) {
    const uint32_t n = x.g.n;
    // list positions: list_buf holds the bucket regions back to back
    uint32_t *base = lists[0] - bucket_off[0];
    uint32_t *valid = ws.get<uint32_t>(kSortSlot + 0, n);
    fill_u32(valid, 0u, n, s);
    for (int b = 0; b < kNumBuckets; ++b)
        if (counts[b])
            hipLaunchKernelGGL(k_sort_positions, dim3(grid_for(counts[b], kBlock * 4)),
                               dim3(kBlock), 0, s, valid, n, bucket_off[b], counts[b]);
    uint64_t *off = ws.get<uint64_t>(kSortSlot + 1, (size_t)n + 1);
    hipLaunchKernelGGL(k_sort_degrees, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, x.g.row,
                       (const uint32_t *)base, n, (const uint32_t *)valid, off);
    set_u64(off + n, 0ull, s);
    scan_u64(off, off, (size_t)n + 1, false,
             ws.get<uint64_t>(kSortSlot + 2, scan_scratch_elems((size_t)n + 1)), s);
    uint64_t total = 0;
    LV_HIP(hipMemcpyAsync(&total, off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    const int bc = bits_for_count((uint64_t)n + 1);
    SortCtx a;
    a.row = x.g.row; a.col = x.g.col; a.wt = x.g.wt; a.comm = x.comm; a.tot = x.tot;
    a.kdeg = x.g.kdeg; a.target = x.target; a.off = off;
    a.keys = ws.get<uint64_t>(kSortSlot + 3, total);
    a.vals = ws.get<uint32_t>(kSortSlot + 4, total);
    a.rbeg = ws.get<uint32_t>(kSortSlot + 5, n);
    a.rend = ws.get<uint32_t>(kSortSlot + 6, n);
    a.bits = bc;
    a.sentinel = (1ull << (2 * bc)) - 1;
    a.v_end = x.v_end;
    a.m2 = x.m2;
    a.tw = x.tw;
    a.highest = x.highest;
    a.bt = x.bt;
    const bool wtd = x.g.wt != nullptr;
    for (int b = 0; b < kNumBuckets; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        a.stats = x.bstats + (size_t)b * kNumStats * kStatSlots;
        const uint32_t lb = (uint32_t)bucket_off[b];
        const uint32_t *L = lists[b];
#define LV_SG(GG, PARTS)                                                                        \
    do {                                                                                        \
        const uint32_t gpb = kBlock / (GG);                                                     \
        const dim3 grid((GG) == kBlock ? cnt * (PARTS) : (cnt + gpb - 1) / gpb);                \
        if (wtd) hipLaunchKernelGGL((k_sort_gather<GG, true>), grid, dim3(kBlock), 0, s, a, L, cnt, lb, (uint32_t)(PARTS)); \
        else hipLaunchKernelGGL((k_sort_gather<GG, false>), grid, dim3(kBlock), 0, s, a, L, cnt, lb, (uint32_t)(PARTS)); \
    } while (0)
        switch (b) {
            case 0: LV_SG(4, 1); break;
            case 1: LV_SG(8, 1); break;
            case 2: LV_SG(16, 1); break;
            case 3: LV_SG(32, 1); break;
            case 4: LV_SG(64, 1); break;
            case 5: LV_SG(64, 1); break;
            case 6: LV_SG(256, 1); break;
            case 7: LV_SG(256, 4); break;
            default: LV_SG(256, kHubParts); break;
        }
#undef LV_SG
    }
    if (total == 0) return;
    uint64_t *k1 = ws.get<uint64_t>(kSortSlot + 7, total);
    uint32_t *v1 = ws.get<uint32_t>(kSortSlot + 8, total);
    radix_sort_pairs(a.keys, a.vals, k1, v1, total, 2 * bc,
                     ws.get<uint32_t>(kSortSlot + 9, radix_hist_elems(total)),
                     ws.get<uint32_t>(kSortSlot + 10, scan_scratch_elems(radix_hist_elems(total))), s);
    // runs of equal (v, c): w(v, c) = difference of the inclusive weight scan
    uint32_t *idx = ws.get<uint32_t>(kSortSlot + 11, total);
    uint64_t *S = ws.get<uint64_t>(kSortSlot + 12, total);
    heads_u64(a.keys, total, idx, s);
    scan_u32(idx, idx, total, true, ws.get<uint32_t>(kSortSlot + 13, scan_scratch_elems(total)), s);
    widen_u32(a.vals, (uint32_t)total, S, s);
    scan_u64(S, S, total, true, ws.get<uint64_t>(kSortSlot + 2, scan_scratch_elems(total)), s);
    uint32_t nu = 0;
    LV_HIP(hipMemcpyAsync(&nu, idx + (total - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    uint64_t *ukey = ws.get<uint64_t>(kSortSlot + 14, nu);
    uint64_t *wsum0 = ws.get<uint64_t>(kSortSlot + 15, nu);
    uint64_t *wsum1 = ws.get<uint64_t>(kSortSlot + 16, nu);
    agg_runs(a.keys, a.vals, idx, S, total, ukey, wsum0, wsum1, s);
    hipLaunchKernelGGL(k_sort_ranges, dim3(grid_for(nu, kBlock * 4)), dim3(kBlock), 0, s,
                       (const uint64_t *)ukey, (uint64_t)nu, bc, n, a.rbeg, a.rend);
    for (int b = 0; b < kNumBuckets; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        a.stats = x.bstats + (size_t)b * kNumStats * kStatSlots;
        const uint32_t *L = lists[b];
#define LV_SD(GG)                                                                               \
    do {                                                                                        \
        const dim3 grid((cnt + kBlock / (GG) - 1) / (kBlock / (GG)));                           \
        if (x.gain == 0) hipLaunchKernelGGL((k_sort_decide<GG, 0>), grid, dim3(kBlock), 0, s, a, L, cnt, ukey, wsum0, wsum1); \
        else hipLaunchKernelGGL((k_sort_decide<GG, 1>), grid, dim3(kBlock), 0, s, a, L, cnt, ukey, wsum0, wsum1); \
    } while (0)
        switch (b) {
            case 0: LV_SD(4); break;
            case 1: LV_SD(8); break;
            case 2: LV_SD(16); break;
            case 3: LV_SD(32); break;
            case 4: LV_SD(64); break;
            case 5: LV_SD(64); break;
            default: LV_SD(256); break;
        }
#undef LV_SD
    }
}

}  // namespace lv
