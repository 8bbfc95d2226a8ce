// lv_kernels.hip -- gfx950 kernels of the Louvain optimisation/aggregation loop.
//
// Best-move pass (reference: phaseOptimizationSort.cu:28-379,
// phaseOptimizationHash.cu:23-162, hashmap.cuh:19-73): vertices are grouped by
// degree; each degree bucket runs a kernel in which a group of G lanes (a
// quarter wave up to a whole workgroup) owns one vertex, accumulates the
// weights towards each neighbour community in an LDS open-addressing table
// (1.5x the degree plus one slot: never more than two thirds full, so every
// probe ends), looks up its own community's weight,
// scores every candidate and reduces to the best one.  Vertices above 4096
// entries are cut into 32 chunks and 32 community hash partitions, each
// partition summed in an LDS table by a workgroup of its own.  The first
// sweep of every level runs at the identity partition and needs no table.
//
// Commit pass (phaseOptimizationUpdate.cu:14-132): movers update the exact
// community totals with integer atomics, add their contribution to the change
// of sigma_in, and mark neighbours left in another community (push form of
// update_changed_kernel; equal to the pull form because the adjacency is
// symmetric).  Finalize pass: comm := target, sum of tot^2, and compaction of
// the marked vertices into per-bucket active lists for the next sweep.
//
// Aggregation (phaseAggregation.cuh:19-175): non-empty flags and inclusive
// scan (renumbering) here; the contraction itself is in lv_aggregate.hip.
#include "lv_kernels.h"

#include "../../include/constants.h"
#include "../../include/device_array.h"
#include "../../include/operators.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <type_traits>

#include "lv_device.h"

namespace lv {

namespace {

// ---------------------------------------------------------------- utilities

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint8_t v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// Start of a sweep: statistics, movers counts and next-list counts to zero.
__global__ void k_sweep_reset(unsigned long long *stats, size_t n_stats, uint32_t *a, uint32_t na,
                              uint32_t *b, uint32_t nb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_stats;
         i += (size_t)gridDim.x * blockDim.x)
        stats[i] = 0ull;
    if (blockIdx.x == 0) {
        if (threadIdx.x < na) a[threadIdx.x] = 0u;
        if (threadIdx.x < nb) b[threadIdx.x] = 0u;
    }
}

__global__ void k_iota_u32(uint32_t *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void k_scatter_u32(uint32_t *dst, const uint32_t *idx, const uint32_t *val, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[idx[i]] = val[i];
}

// Plain streaming copy (16 bytes per lane per step): the bandwidth probe and
// the known-size reference for hardware traffic counters.
__global__ __launch_bounds__(kBlock) void k_stream_copy(const uint4 *src, uint4 *dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

constexpr int kProbeIlp = 4;

// Random-gather probe: each lane streams kProbeIlp column ids and gathers the
// element arr[col[j]] of each (4 or 8 bytes), the one dependent access of
// the best-move pass without its table work: the measured ceiling of that
// access pattern on this graph.  The sum keeps the loads live.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_probe(const uint32_t *__restrict__ col,
                                                         uint64_t E, const T *__restrict__ arr,
                                                         unsigned long long *out) {
    unsigned long long acc = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j0 < E;
         j0 += stride * kProbeIlp) {
        uint32_t x[kProbeIlp];
#pragma unroll
        for (int u = 0; u < kProbeIlp; ++u) {
            const uint64_t j = j0 + (uint64_t)u * stride;
            x[u] = j < E ? stream_load(&col[j]) : 0u;
        }
        T e[kProbeIlp];
#pragma unroll
        for (int u = 0; u < kProbeIlp; ++u) e[u] = arr[x[u]];
#pragma unroll
        for (int u = 0; u < kProbeIlp; ++u) acc += (unsigned long long)e[u];
    }
    out[blockIdx.x * (uint64_t)blockDim.x + threadIdx.x] = acc;
}

// Non-self entries of every vertex (rows of contracted levels are unsorted).
__global__ void k_nonself_counts(const uint64_t *row, const uint32_t *col, uint32_t n,
                                 uint32_t *nsel) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        uint32_t c = 0;
        for (uint64_t j = row[v]; j < row[v + 1]; ++j) c += col[j] != (uint32_t)v;
        nsel[v] = c;
    }
}

// Selected entries per reference chunk for the listed vertices; chunk k holds
// vertices [cv[k], cv[k+1]).
__global__ __launch_bounds__(kBlock) void k_chunk_selected(const uint32_t *list, uint32_t count,
                                                          const uint32_t *nsel, const uint32_t *cv,
                                                          uint32_t nc, unsigned long long *csel) {
    __shared__ unsigned long long h[kMaxChunks];
    for (uint32_t i = threadIdx.x; i < nc; i += kBlock) h[i] = 0;
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t v = list[i];
        uint32_t lo = 0, hi = nc;   // largest k with cv[k] <= v
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cv[mid] <= v) lo = mid; else hi = mid;
        }
        const uint32_t c = nsel[v];
        if (c) atomicAdd(&h[lo], (unsigned long long)c);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nc; i += kBlock)
        if (h[i]) atomicAdd(&csel[i], h[i]);
}

__global__ void k_gather_degrees(const uint64_t *row, const uint32_t *ids, uint32_t n,
                                 uint64_t *deg) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        deg[i] = row[ids[i] + 1] - row[ids[i]];
}

// COO form of a CSR: the source of every entry, and every row's length.
__global__ void k_coo_sources(const uint64_t *row, uint32_t n, uint32_t *src, uint32_t *deg) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        const uint64_t b = row[v], e = row[v + 1];
        deg[v] = (uint32_t)(e - b);
        for (uint64_t j = b; j < e; ++j) src[j] = (uint32_t)v;
    }
}

__global__ void k_community_totals(const uint32_t *comm, const uint32_t *kdeg, uint32_t n,
                                   uint32_t *tot) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x)
        atomicAdd(&tot[comm[v]], kdeg[v]);
}


// ------------------------------------------------------------------- scan

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename T>
__device__ inline T wave_incl_scan(T x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T y;
        if constexpr (sizeof(T) == 8) {
            const int lo = __shfl_up((int)(uint32_t)(uint64_t)x, off);
            const int hi = __shfl_up((int)(uint32_t)((uint64_t)x >> 32), off);
            y = (T)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
        } else {
            y = (T)__shfl_up((int)x, off);
        }
        if (lane >= off) x += y;
    }
    return x;
}

// Exclusive block scan of one value per thread; returns the block total in *total.
template <typename T>
__device__ inline T block_excl_scan(T x, T *total) {
    __shared__ T wsum[kBlock / 64];
    const T inc = wave_incl_scan<T>(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
    for (int i = 0; i < kBlock / 64; ++i) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - x;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_tile(const T *in, T *out, size_t n, T *block_sums,
                                                      int inclusive) {
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T run = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const size_t e = base + i;
        v[i] = e < n ? in[e] : (T)0;
        run += v[i];
    }
    T total;
    T off = block_excl_scan<T>(run, &total);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const size_t e = base + i;
        const T x = v[i];
        if (e < n) out[e] = inclusive ? off + x : off;
        off += x;
    }
    if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_add(T *out, size_t n, const T *block_off) {
    const T add = block_off[blockIdx.x];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kBlock) {
        const size_t e = base + i;
        if (e < n) out[e] += add;
    }
}

template <typename T>
void scan_impl(const T *in, T *out, size_t n, bool inclusive, T *scratch, hipStream_t s) {
    if (n == 0) return;
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tile<T>, dim3(1), dim3(kBlock), 0, s, in, out, n, (T *)nullptr,
                           inclusive ? 1 : 0);
        return;
    }
    T *sums = scratch;
    T *rest = scratch + nb;
    hipLaunchKernelGGL(k_scan_tile<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, in, out, n, sums,
                       inclusive ? 1 : 0);
    scan_impl<T>(sums, sums, nb, false, rest, s);
    hipLaunchKernelGGL(k_scan_add<T>, dim3((unsigned)nb), dim3(kBlock), 0, s, out, n,
                       (const T *)sums);
}

// ------------------------------------------------------------- radix sort

constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;

__global__ __launch_bounds__(kBlock) void k_rs_upsweep(const uint64_t *keys, size_t n, int shift,
                                                       uint32_t *hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
    for (int i = 0; i < kRsItems; ++i) {
        const size_t e = base + (size_t)i * kBlock + threadIdx.x;
        if (e < n) atomicAdd(&h[(uint32_t)(keys[e] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: tile element e = i*256 + tid keeps its order among equal digits.
__global__ __launch_bounds__(kBlock) void k_rs_downsweep(const uint64_t *kin, const uint32_t *vin,
                                                         uint64_t *kout, uint32_t *vout, size_t n,
                                                         int shift, const uint32_t *hist_off,
                                                         uint32_t nblocks) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_w[kBlock / 64][256];
    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    s_base[tid] = hist_off[(size_t)tid * nblocks + blockIdx.x];
    s_run[tid] = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s_w[w][tid] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
    const unsigned long long lt = lanemask_lt();
    for (int i = 0; i < kRsItems; ++i) {
        const size_t e = base + (size_t)i * kBlock + tid;
        const bool valid = e < n;
        uint64_t key = 0;
        uint32_t val = 0;
        uint32_t d = 0;
        if (valid) {
            key = kin[e];
            val = vin[e];
            d = (uint32_t)(key >> shift) & 0xFFu;
        }
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t lower = (uint32_t)__popcll(peers & lt);
        if (valid && lower == 0) s_w[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t r = s_run[tid];
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) {
                const uint32_t c = s_w[w][tid];
                s_w[w][tid] = r;
                r += c;
            }
            s_run[tid] = r;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = s_base[d] + s_w[wave][d] + lower;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) s_w[w][tid] = 0;
        __syncthreads();
    }
}

// ------------------------------------------------------------ level setup

__global__ void k_bucket_ids(const uint64_t *row, uint32_t n, uint8_t *bucket, BucketBounds bounds) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x)
        bucket[v] = bucket_of_degree(row[v + 1] - row[v], bounds);
}


__global__ __launch_bounds__(kBlock) void k_bucket_counts(const uint8_t *bucket, uint32_t n,
                                                          uint32_t *counts) {
    __shared__ uint32_t h[kNumBuckets];
    if (threadIdx.x < kNumBuckets) h[threadIdx.x] = 0;
    __syncthreads();
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x)
        atomicAdd(&h[bucket[v]], 1u);
    __syncthreads();
    if (threadIdx.x < kNumBuckets && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

// Every vertex appended to the list of its bucket; each workgroup takes
// kListTiles consecutive tiles of kBlock vertices and reserves its run of
// every list with one atomic per bucket.
constexpr int kListTiles = 16;

__global__ __launch_bounds__(kBlock) void k_build_lists_all(const uint8_t *bucket, uint32_t n,
                                                            ListPtrs lists, uint32_t *counts) {
    __shared__ uint32_t s_cnt[kNumBuckets];
    __shared__ uint32_t s_base[kNumBuckets];
    if (threadIdx.x < kNumBuckets) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t v0 = blockIdx.x * (uint32_t)(kBlock * kListTiles);
    uint32_t rank[kListTiles];
    uint8_t bk[kListTiles];
#pragma unroll
    for (int i = 0; i < kListTiles; ++i) {
        const uint32_t v = v0 + (uint32_t)i * kBlock + threadIdx.x;
        bk[i] = v < n ? bucket[v] : 0;
        rank[i] = v < n ? atomicAdd(&s_cnt[bk[i]], 1u) : LV_EMPTY;
    }
    __syncthreads();
    if (threadIdx.x < kNumBuckets) {
        const uint32_t c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&counts[threadIdx.x], c) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kListTiles; ++i)
        if (rank[i] != LV_EMPTY)
            lists.p[bk[i]][s_base[bk[i]] + rank[i]] = v0 + (uint32_t)i * kBlock + threadIdx.x;
}

// First entry index handled by a thread chunk, and its source vertex.
__device__ inline uint32_t src_of_entry(const uint64_t *row, uint32_t n, uint64_t j) {
    uint32_t lo = 0, hi = n;   // largest v with row[v] <= j
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (row[mid] <= j) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int kChunk = 16;   // entries per thread in entry-parallel kernels
constexpr int kIlp = 4;      // entries per lane with loads in flight together

// sigma_in = sum of w over entries with comm[src]==comm[dst]; sum of tot^2.
__global__ __launch_bounds__(kBlock) void k_modularity_parts(LevelView g, const uint32_t *comm,
                                                             const uint32_t *tot,
                                                             unsigned long long *stats) {
    uint64_t sig = 0, sq = 0;
    const uint64_t nchunks = (g.E + kChunk - 1) / kChunk;
    for (uint64_t ch = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; ch < nchunks;
         ch += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j0 = ch * kChunk;
        const uint64_t j1 = j0 + kChunk < g.E ? j0 + kChunk : g.E;
        uint32_t v = src_of_entry(g.row, g.n, j0);
        uint64_t vend = g.row[v + 1];
        uint32_t cv = comm[v];
        for (uint64_t j = j0; j < j1; ++j) {
            while (j >= vend) { ++v; vend = g.row[v + 1]; cv = comm[v]; }
            if (internal_entry(cv, comm[g.col[j]])) sig += g.wt ? g.wt[j] : 1u;
        }
    }
    for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < g.n;
         c += (uint64_t)gridDim.x * blockDim.x)
        sq += square_u64(tot[c]);
    block_add_stats(stats, ST_DSIGMA, sig, ST_SUMSQ, sq, -1, 0);
}

// Identity partition: sigma_in = weight of self-loop entries (rows are sorted
// by destination, so v's own entries are found by binary search), sum of k^2.
__global__ __launch_bounds__(kBlock) void k_identity_parts(LevelView g,
                                                           unsigned long long *stats) {
    uint64_t sig = 0, sq = 0;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < g.n;
         v += (size_t)gridDim.x * blockDim.x) {
        uint64_t lo = g.row[v], hi = g.row[v + 1];
        const uint64_t end = hi;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (g.col[mid] < v) lo = mid + 1; else hi = mid;
        }
        for (uint64_t j = lo; j < end && g.col[j] == v; ++j) sig += g.wt ? g.wt[j] : 1u;
        sq += square_u64(g.kdeg[v]);
    }
    block_add_stats(stats, ST_DSIGMA, sig, ST_SUMSQ, sq, -1, 0);
}

// ---------------------------------------------------------- best-move pass

struct MoveCtx {
    BatchView bt;
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *wt;
    const uint32_t *kdeg;
    const uint32_t *comm;
    const uint32_t *tot;
    const unsigned long long *ctot;   // comm | tot[comm] << 32 of every vertex (hash sweeps)
    uint32_t *target;
    unsigned long long *stats;
    int64_t m2;
    double tw;
    uint32_t v_end;
    bool highest;
};

// The words of one listed vertex a best-move table kernel needs before its
// row: loaded together (their latencies overlap), for the vertex at hand or,
// in the persistent kernels, for the next one while this one is scored.
struct MoveWords {
    uint32_t v = 0, own = 0, kv = 0, tot_own = 0;
    uint64_t beg = 0, end = 0;
    int64_t m2 = 0;
    double tw = 0;
    bool live = false;
};

// A load the compiler may issue through the scalar unit when its address is
// uniform: arrays a kernel only reads (row bounds, degrees, comm|tot words)
// are read as constant memory, so a vertex's words land in scalar registers
// and stay there while the vector registers hold its row.
template <typename T>
__device__ inline T const_load(const T *p) {
    return *reinterpret_cast<const __attribute__((address_space(4))) T *>(
        reinterpret_cast<uintptr_t>(p));
}

template <bool PK, bool SCALAR = false>
__device__ inline MoveWords move_words(const MoveCtx &a, uint32_t v, bool valid) {
    auto ld = [](const auto *p) { return SCALAR ? const_load(p) : *p; };
    MoveWords w;
    w.v = v;
    w.m2 = a.m2;
    w.tw = a.tw;
    w.live = valid && v < a.v_end && vertex_graph(a.bt, v, a.m2, a.tw, w.m2, w.tw);
    // (scalar: loaded whether live or not, from a vertex that exists, so
    // that no wait is needed to merge the two cases)
    if (SCALAR || w.live) {
        if constexpr (PK) {
            const unsigned long long p = ld(&a.ctot[v]);
            w.own = (uint32_t)p;
            w.tot_own = (uint32_t)(p >> 32);
        } else {
            w.own = ld(&a.comm[v]);
        }
        w.kv = ld(&a.kdeg[v]);
        w.beg = ld(&a.row[v]);
        w.end = ld(&a.row[v + 1]);
        if constexpr (!PK) w.tot_own = ld(&a.tot[w.own]);
    }
    return w;
}

// Group-per-vertex with an LDS table of T slots (T >= table_slots of the
// bucket's maximum degree); a vertex uses the first table_slots(degree).
// PK: a neighbour's community and that community's total come in one 8-byte
// gather of ctot (comm | tot << 32), and the first lane to insert a community
// keeps its total in a third LDS array; the candidates are then scored from
// LDS alone.  Without PK the totals are gathered once per candidate (one more
// dependent random access per distinct community, for 4 bytes less LDS per
// slot).
// G == kBlock (one workgroup per vertex) runs persistent: a workgroup walks
// the list with a stride of the grid, the next vertex's list entry is loaded
// when a vertex starts and its words when its inserts are done, so both
// latencies pass under this vertex's barrier and scoring, and the statistics
// are reduced once per workgroup.  The grid is a few rounds of the
// workgroups the device holds at once (kPersistRounds).
template <int G, int T, int GAIN, bool WEIGHTED, bool PK>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(G == kBlock && PK && GAIN == 0 ? 8 : 1))) void k_move_table(MoveCtx a, const uint32_t *__restrict__ list,
                                                       uint32_t count) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    constexpr int GPB = kBlock / G;
    constexpr bool PERSIST = G == kBlock;
    constexpr int SI = PERSIST && PK ? 1 : kIlp;
    // a slot is (key, summed weight), adjacent: one 8-byte LDS access reads
    // or clears both
    __shared__ uint64_t s_kv[GPB * T];
    __shared__ uint32_t s_tot[PK ? GPB * T : 1];
    __shared__ S r_s[kBlock / 64];
    __shared__ uint32_t r_c[kBlock / 64];
    __shared__ uint32_t r_n[kBlock / 64];
    const int lane = threadIdx.x & (G - 1);
    const int gi = (int)group_index<G>();
    uint64_t *KV = s_kv + gi * T;
    uint32_t *const KW = reinterpret_cast<uint32_t *>(KV);   // 2h: key, 2h + 1: weight
    constexpr uint64_t kEmptySlot = (uint64_t)LV_EMPTY;       // key LV_EMPTY, weight 0
    uint32_t *TT = s_tot + (PK ? gi * T : 0);
    uint32_t idx = PERSIST ? blockIdx.x : blockIdx.x * GPB + gi;
    MoveWords nw = move_words<PK, PERSIST>(a, idx < count ? list[idx] : 0u, idx < count);
    // persistent: the list entry of the vertex after the next one, loaded a
    // vertex ahead, so that the next vertex's words are requested when a
    // vertex starts and arrive under its inserts
    uint32_t nv = 0u;
    if constexpr (PERSIST) {
        const uint32_t n1 = idx + gridDim.x;
        nv = n1 > idx && n1 < count ? list[n1] : 0u;
    }
    uint32_t acc_self = 0, acc_cand = 0, acc_cons = 0, acc_tw = 0;
    // persistent: thread 0 counts its decisions in LDS (two registers fewer
    // in every lane)
    __shared__ uint32_t s_cons, s_tw;
    if (PERSIST && threadIdx.x == 0) { s_cons = 0u; s_tw = 0u; }
    // persistent: the first kIlp columns per lane of the next vertex's row,
    // loaded while this vertex is scored (its weights load in its first step)
    uint32_t px[kIlp];
    auto first_loads = [&](const MoveWords &w) {
#pragma unroll
        for (int u = 0; u < kIlp; ++u) {
            const uint64_t j = w.beg + (uint64_t)lane + (uint64_t)u * G;
            px[u] = w.live && j < w.end ? stream_load(&a.col[j]) : w.v;
        }
    };
    constexpr bool PREFETCH = PERSIST;
    // ... and, packed, the (comm, total) words of those columns, gathered
    // once this vertex's candidates are scored, before its reduction
    // (unpacked: their communities)
    constexpr bool PGATHER = PREFETCH;
    unsigned long long pg[PGATHER ? kIlp : 1];
    auto first_gathers = [&]() {
        if constexpr (PGATHER) {
#pragma unroll
            for (int u = 0; u < kIlp; ++u) pg[u] = PK ? a.ctot[px[u]] : a.comm[px[u]];
        }
    };
    if constexpr (PREFETCH) first_loads(nw);
    first_gathers();
    // FUSED (persistent, exact gains): the table is cleared once here, then
    // by the scoring loop as it reads each slot (every slot belongs to one
    // lane there), so a vertex starts on a clean table with no clearing loop
    // and no barrier of its own; the own community's weight, needed only by
    // the final test, is posted by the lane that meets it
    constexpr bool FUSED = PERSIST && GAIN == 0;
    __shared__ uint32_t s_selfw;
    if constexpr (FUSED) {
        for (uint32_t s = lane; s < (uint32_t)T; s += G) KV[s] = kEmptySlot;
        if (threadIdx.x == 0) s_selfw = 0u;
        __syncthreads();
    }
    for (;;) {
        const MoveWords cw = nw;
        const uint32_t v = cw.v, own = cw.own, kv = cw.kv, tot_own = cw.tot_own;
        const uint64_t beg = cw.beg, end = cw.end;
        const bool live = cw.live;
        const uint32_t nidx = idx + gridDim.x;   // PERSIST only
        const bool nvalid = PERSIST && nidx > idx && nidx < count;
        if constexpr (PERSIST) {
            nw = move_words<PK, true>(a, nv, nvalid);
            const uint32_t n2 = nidx + gridDim.x;
            nv = nvalid && n2 > nidx && n2 < count ? list[n2] : 0u;
        }
        // the vertex's table: table_slots(degree) slots, at most T (which
        // exceeds every degree of the bucket)
        const uint32_t tv = live ? min(table_slots(end - beg), (uint32_t)T) : 4u;
        if constexpr (!FUSED) {
            for (uint32_t s = lane; s < tv; s += G) KV[s] = kEmptySlot;
            // a vertex's table is its group's alone: groups inside one wave
            // wait for no other wave of the workgroup
            if constexpr (PERSIST) lds_barrier();
            else group_barrier<G>();
        }
        uint32_t nself = 0;
        // I entries per lane in flight: loads first, then gathers, then
        // inserts; `first`: the row's first step, whose columns the persistent
        // kernels prefetched
        auto step = [&](auto ilp, uint64_t jb, bool first) {
            constexpr int I = decltype(ilp)::value;
            uint32_t xs[I], ws[I], cs[I], ts[I];
            if (PREFETCH && first) {
#pragma unroll
                for (int u = 0; u < I; ++u) {
                    const uint64_t j = jb + (uint64_t)u * G;
                    xs[u] = px[u];
                    if constexpr (WEIGHTED) ws[u] = stream_load(&a.wt[j < end ? j : end - 1]);
                    else ws[u] = 1u;
                }
            } else {
#pragma unroll
                for (int u = 0; u < I; ++u) {
                    const uint64_t j = jb + (uint64_t)u * G;
                    if constexpr (G >= 64) {
                        // past the row's end a lane loads the row's last
                        // entry (no branch; the lanes of a group share that
                        // address) and takes the vertex itself, a skipped entry
                        const uint64_t jc = j < end ? j : end - 1;
                        const uint32_t x = stream_load(&a.col[jc]);
                        xs[u] = j < end ? x : v;
                        if constexpr (WEIGHTED) ws[u] = stream_load(&a.wt[jc]);
                        else ws[u] = 1u;
                    } else {
                        xs[u] = j < end ? stream_load(&a.col[j]) : v;
                        ws[u] = (WEIGHTED && j < end) ? stream_load(&a.wt[j]) : 1u;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < I; ++u) {
                // groups of a wave or more: every xs[u] is a vertex id and
                // the words are gathered without a branch (those of
                // skipped entries are not used; the group's skipped lanes
                // share one address)
                const bool g = G >= 64 || xs[u] != v;
                if constexpr (PK) {
                    const unsigned long long p =
                        PGATHER && first ? pg[PGATHER ? u : 0] : g ? a.ctot[xs[u]] : 0ull;
                    cs[u] = (uint32_t)p;
                    ts[u] = (uint32_t)(p >> 32);
                } else {
                    cs[u] = PGATHER && first ? (uint32_t)pg[PGATHER ? u : 0] : g ? a.comm[xs[u]] : 0u;
                    ts[u] = 0u;
                }
            }
            // the first probe of all I entries is issued back to back (one
            // LDS round trip instead of I); a lane's own entries are still
            // inserted in order, so an entry sees the keys its predecessors
            // placed, and the rare occupied first slots probe on afterwards
            uint32_t hs[I], pv[I];
#pragma unroll
            for (int u = 0; u < I; ++u) {
                hs[u] = slot_hash_mod(cs[u], tv);
                pv[u] = xs[u] != v ? atomicCAS(&KW[2 * hs[u]], LV_EMPTY, cs[u]) : LV_EMPTY;
            }
#pragma unroll
            for (int u = 0; u < I; ++u) {
                if (xs[u] == v) continue;
                ++nself;
                const uint32_t c = cs[u];
                uint32_t h = hs[u], prev = pv[u];
                while (prev != LV_EMPTY && prev != c) {
                    h = h + 1 == tv ? 0u : h + 1;
                    prev = atomicCAS(&KW[2 * h], LV_EMPTY, c);
                }
                if (PK && prev == LV_EMPTY) TT[h] = ts[u];
                atomicAdd(&KW[2 * h + 1], ws[u]);
            }
        };
        if (live) {
            if (PERSIST && end - beg <= (uint64_t)G * (kIlp / 2)) {
                // short rows of a persistent kernel's bucket (group-uniform
                // test): half the entries per lane, no masked half step
                step(std::integral_constant<int, kIlp / 2>{}, beg + lane, true);
            } else if constexpr (PERSIST) {
                // the first step outside the loop: the loop's entry would wait
                // for the next vertex's scalar words, which the first step's
                // gathers wait for anyway
                step(std::integral_constant<int, kIlp>{}, beg + lane, true);
                for (uint64_t jb = beg + lane + (uint64_t)G * kIlp; jb < end; jb += (uint64_t)G * kIlp)
                    step(std::integral_constant<int, kIlp>{}, jb, false);
            } else {
                for (uint64_t jb = beg + lane; jb < end; jb += (uint64_t)G * kIlp)
                    step(std::integral_constant<int, kIlp>{}, jb, jb == beg + lane);
            }
        }
        if constexpr (PERSIST) lds_barrier();
        else group_barrier<G>();
        if constexpr (PREFETCH) first_loads(nw);
        S best = SC::none();
        uint32_t best_c = a.highest ? 0u : LV_EMPTY;
        uint32_t self = 0;
        if (live) {
            if constexpr (!FUSED) {
                uint32_t h = slot_hash_mod(own, tv);
                for (;;) {
                    const uint64_t kv = KV[h];
                    const uint32_t c = (uint32_t)kv;
                    if (c == own) { self = (uint32_t)(kv >> 32); break; }
                    if (c == LV_EMPTY) break;
                    h = h + 1 == tv ? 0u : h + 1;
                }
            }
            // branch free: slots past the table read slot 0, and a slot that
            // holds no other community is scored but never taken
            for (uint32_t sb = lane; sb < tv; sb += (uint32_t)G * SI) {
                uint32_t cs[SI], Ws[SI], ts[SI];
                bool ok[SI];
#pragma unroll
                for (int u = 0; u < SI; ++u) {
                    const uint32_t s2 = sb + (uint32_t)u * G;
                    const uint32_t si = s2 < tv ? s2 : 0u;
                    const uint64_t kv = KV[si];
                    cs[u] = (uint32_t)kv;
                    Ws[u] = (uint32_t)(kv >> 32);
                    ok[u] = (s2 < tv) & (cs[u] != own) & (cs[u] != LV_EMPTY);
                    if constexpr (FUSED) {
                        if (s2 < tv) {
                            if (cs[u] == own) s_selfw = Ws[u];
                            KV[s2] = kEmptySlot;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < SI; ++u) {
                    const uint32_t s2 = sb + (uint32_t)u * G;
                    if constexpr (PK) ts[u] = TT[s2 < tv ? s2 : 0u];
                    else ts[u] = ok[u] ? a.tot[cs[u]] : 0u;
                }
#pragma unroll
                for (int u = 0; u < SI; ++u) {
                    acc_cand += ok[u];
                    const S sc = SC::of(Ws[u], self, kv, tot_own, ts[u], cw.m2, cw.tw);
                    const bool b = ok[u] & better_candidate(sc, cs[u], best, best_c, a.highest);
                    best = b ? sc : best;
                    best_c = b ? cs[u] : best_c;
                }
            }
        }
        acc_self += nself;
        first_gathers();
        if constexpr (G <= 64) {
            group_best<G, S>(best, best_c, a.highest);
            const bool gself = group_any<G>(nself > 0);
            if (live && lane == 0) {
                acc_cons += gself;
                if (best != SC::none() && SC::positive(best, self, kv, tot_own, cw.m2)) {
                    a.target[v] = best_c;
                    ++acc_tw;
                }
            }
        } else {
            group_best<64, S>(best, best_c, a.highest);
            const uint32_t wself = group_any<64>(nself > 0);
            const int w = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; r_n[w] = wself; }
            if constexpr (PERSIST) lds_barrier();
            else __syncthreads();
            if (threadIdx.x == 0 && live) {
                uint32_t tn = 0;
                for (int i = 0; i < kBlock / 64; ++i) {
                    tn += r_n[i];
                    if (better_candidate(r_s[i], r_c[i], best, best_c, a.highest)) {
                        best = r_s[i];
                        best_c = r_c[i];
                    }
                }
                if constexpr (FUSED) {
                    self = s_selfw;
                    s_selfw = 0u;
                }
                s_cons += tn > 0;
                if (best != SC::none() && SC::positive(best, self, kv, tot_own, cw.m2)) {
                    a.target[v] = best_c;
                    ++s_tw;
                }
            }
        }
        // uniform over the workgroup: idx and the grid are the same for all lanes
        if (!nvalid) break;
        idx = nidx;
    }
    if (PERSIST && threadIdx.x == 0) { acc_cons = s_cons; acc_tw = s_tw; }
    block_add_stats(a.stats, ST_E_ACTIVE, acc_self, ST_CANDIDATES, acc_cand, ST_CONSIDERED,
                    acc_cons, ST_TWRITES, acc_tw);
}

// First sweep of a level (identity partition, phaseOptimizationSort.cu:28-158):
// every non-self entry is its own candidate community, self = 0, and a move
// also needs v < c (phaseOptimizationUpdate.cu:99).  Ties: lowest id.
template <int G, int GAIN, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_move_fast(MoveCtx a, const uint32_t *__restrict__ list,
                                                      uint32_t count) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    constexpr int GPB = kBlock / G;
    const int lane = threadIdx.x & (G - 1);
    const int gi = (int)group_index<G>();
    const uint32_t idx = blockIdx.x * GPB + gi;
    const bool valid = idx < count;
    const uint32_t v = valid ? list[idx] : 0u;
    int64_t m2 = a.m2;
    double tw = a.tw;
    const bool live = valid && v < a.v_end && vertex_graph(a.bt, v, a.m2, a.tw, m2, tw);
    S best = SC::none();
    uint32_t best_c = LV_EMPTY;
    uint32_t nself = 0;
    uint32_t kv = 0, tot_own = 0;
    if (live) {
        kv = a.kdeg[v];
        tot_own = a.tot[v];
        const uint64_t beg = a.row[v], end = a.row[v + 1];
        // candidates tested without branches
        for (uint64_t jb = beg + lane; jb < end; jb += (uint64_t)G * kIlp) {
            uint32_t xs[kIlp], ws[kIlp], ts[kIlp];
#pragma unroll
            for (int u = 0; u < kIlp; ++u) {
                const uint64_t j = jb + (uint64_t)u * G;
                xs[u] = j < end ? stream_load(&a.col[j]) : v;
                ws[u] = (WEIGHTED && j < end) ? stream_load(&a.wt[j]) : 1u;
            }
#pragma unroll
            for (int u = 0; u < kIlp; ++u) ts[u] = xs[u] != v ? a.tot[xs[u]] : 0u;
#pragma unroll
            for (int u = 0; u < kIlp; ++u) {
                const bool ok = xs[u] != v;
                nself += ok;
                const S sc = SC::of(ws[u], 0u, kv, tot_own, ts[u], m2, tw);
                const bool b = ok & better_candidate(sc, xs[u], best, best_c, false);
                best = b ? sc : best;
                best_c = b ? xs[u] : best_c;
            }
        }
    }
    uint32_t considered = 0, twrite = 0;
    if constexpr (G <= 64) {
        group_best<G, S>(best, best_c, false);
        const bool gself = group_any<G>(nself > 0);
        if (live && lane == 0) {
            considered = gself;
            if (best != SC::none() && SC::positive(best, 0u, kv, tot_own, m2) && v < best_c) {
                a.target[v] = best_c;
                twrite = 1;
            }
        }
    } else {
        __shared__ S r_s[kBlock / 64];
        __shared__ uint32_t r_c[kBlock / 64];
        __shared__ uint32_t r_n[kBlock / 64];
        group_best<64, S>(best, best_c, false);
        const uint32_t wself = group_any<64>(nself > 0);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; r_n[w] = wself; }
        __syncthreads();
        if (threadIdx.x == 0 && live) {
            uint32_t tn = 0;
            for (int i = 0; i < kBlock / 64; ++i) {
                tn += r_n[i];
                if (better_candidate(r_s[i], r_c[i], best, best_c, false)) {
                    best = r_s[i];
                    best_c = r_c[i];
                }
            }
            considered = tn > 0;
            if (best != SC::none() && SC::positive(best, 0u, kv, tot_own, m2) && v < best_c) {
                a.target[v] = best_c;
                twrite = 1;
            }
        }
    }
    block_add_stats(a.stats, ST_E_ACTIVE, nself, ST_CANDIDATES, nself, ST_CONSIDERED, considered,
                    ST_TWRITES, twrite);
}

// ---- hubs: kHubParts workgroups per vertex, partitioned by community hash ----
//
// A hub row is cut into kHubChunks contiguous chunks (one workgroup each) and
// its neighbour communities into kHubParts partitions by the top bits of a
// bijective mix of the community id.  k_hub_count gathers the community of
// every entry, sums the weight towards the hub's own community, folds
// repeats of the communities it caches, writes the chunk's (community,
// weight) pairs packed at the start of the chunk's range of the stage buffer,
// reserves each (chunk, partition) run and posts the hub's reduce groups (one
// per kHubGroupEntries entries, each a run of consecutive partitions);
// k_hub_scatter moves a chunk's pairs to partition order; k_hub_reduce, a persistent grid over the posted groups, sums a
// group's weights per community in an LDS table and scores them; k_hub_decide
// reduces the groups' bests.  Every entry costs streaming reads and writes and
// LDS atomics only; hub rows see almost only distinct communities, so a
// global table would take one random atomic per entry.

struct HubCtx {
    MoveCtx m;
    HubTables h;
};

constexpr int kHubPartBits = ilog2(kHubParts);
constexpr int kHubSubBits = 32 - kHubPartBits;            // mix bits below the partition
constexpr uint64_t kHubGroupEntries = 4096;                // entries per reduce group
// Grid of the persistent k_hub_reduce: about eight times the workgroups
// resident at once.  The items are uneven, and a grid that the device runs in
// several rounds balances them better than one that keeps every workgroup
// resident (measured: 2,048 and the resident count were slower).
constexpr uint32_t kHubReduceBlocks = 8192;
constexpr int kHubIlp = 16;                                // pair loads / tot gathers in flight per lane
static_assert(kHubParts == 32, "items pack the group in 5 bits");
static_assert(kHubChunks >= 1 && kHubChunks <= kHubParts, "fast-sweep chunks use the group slots");

__device__ inline uint32_t ilog2_u32(uint32_t x) { return 31u - (uint32_t)__builtin_clz(x); }

// Reduce groups of a hub row of d entries: a power of two <= kHubParts, one
// per kHubGroupEntries; group g takes kHubParts / groups consecutive partitions.
__device__ inline uint32_t hub_groups(uint64_t d) {
    uint32_t g = 1;
    while (g < (uint32_t)kHubParts && (uint64_t)g * kHubGroupEntries < d) g <<= 1;
    return g;
}

__device__ inline void hub_chunk(uint64_t beg, uint64_t end, uint32_t q, uint64_t &cb,
                                 uint64_t &ce) {
    const uint64_t d = end - beg;
    const uint64_t per = (d + kHubChunks - 1) / kHubChunks;
    cb = beg + min(d, per * q);
    ce = beg + min(d, per * (q + 1));
}

// The hub of workgroup block/kHubParts when it sweeps: listed, in range and of
// a live graph.
__device__ inline bool hub_runs(const HubCtx &a, const uint32_t *list, uint32_t count,
                                uint32_t hi, uint32_t &v, int64_t &m2, double &tw) {
    if (hi >= count) return false;
    v = list[hi];
    return v < a.m.v_end && vertex_graph(a.m.bt, v, a.m.m2, a.m.tw, m2, tw);
}

// Communities cached per chunk workgroup: a repeat adds its weight in LDS, so
// a row whose neighbours share few communities (hubs of contracted levels)
// passes few pairs on to the scatter and the reduce.
constexpr uint32_t kHubCache = 1024;

template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_hub_count(HubCtx a, const uint32_t *__restrict__ list,
                                                      uint32_t count) {
    __shared__ uint32_t hist[kHubParts];
    __shared__ uint64_t r_self[kBlock / 64];
    __shared__ uint32_t c_key[kHubCache];
    __shared__ uint32_t c_val[kHubCache];
    __shared__ uint32_t s_n;   // pairs written so far
    const uint32_t q = blockIdx.x % kHubChunks;
    if (threadIdx.x < kHubParts) hist[threadIdx.x] = 0;
    for (uint32_t i = threadIdx.x; i < kHubCache; i += kBlock) { c_key[i] = LV_EMPTY; c_val[i] = 0u; }
    if (threadIdx.x == 0) s_n = 0u;
    __syncthreads();
    uint32_t v = 0, slot = 0;
    int64_t m2;
    double tw;
    const bool run = hub_runs(a, list, count, blockIdx.x / kHubChunks, v, m2, tw);
    uint64_t nself = 0, selfw = 0;
    unsigned long long *out = nullptr;
    if (run) {
        slot = a.h.slot_of[v];
        const uint32_t own = a.m.comm[v];
        const uint64_t beg = a.m.row[v], end = a.m.row[v + 1];
        uint64_t cb, ce;
        hub_chunk(beg, end, q, cb, ce);
        // the chunk's range of the stage buffer: at most one pair per entry
        out = a.h.stage + a.h.ent_off[slot] + (cb - beg);
        for (uint64_t jb = cb + threadIdx.x; jb < ce; jb += (uint64_t)kBlock * kIlp) {
            uint32_t xs[kIlp], cs[kIlp], ws[kIlp];
#pragma unroll
            for (int u = 0; u < kIlp; ++u) {
                const uint64_t j = jb + (uint64_t)u * kBlock;
                xs[u] = j < ce ? stream_load(&a.m.col[j]) : v;
                ws[u] = (WEIGHTED && j < ce) ? stream_load(&a.m.wt[j]) : 1u;
            }
#pragma unroll
            for (int u = 0; u < kIlp; ++u) cs[u] = xs[u] != v ? a.m.comm[xs[u]] : LV_EMPTY;
#pragma unroll
            for (int u = 0; u < kIlp; ++u) {
                const uint32_t c = cs[u];
                bool emit = false;
                if (c != LV_EMPTY) {
                    ++nself;
                    if (c == own) {
                        selfw += ws[u];
                    } else {
                        const uint32_t k = hub_mix(c);
                        const uint32_t cs_ = k & (kHubCache - 1);
                        uint32_t cur = c_key[cs_];
                        if (cur == LV_EMPTY) cur = atomicCAS(&c_key[cs_], LV_EMPTY, c);
                        if (cur == LV_EMPTY || cur == c) {   // cached here, or a repeat
                            atomicAdd(&c_val[cs_], ws[u]);
                        } else {
                            emit = true;
                            atomicAdd(&hist[k >> kHubSubBits], 1u);
                        }
                    }
                }
                wave_append(emit, ((unsigned long long)c << 32) | ws[u], &s_n, out);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHubCache; i += kBlock) {   // uniform trip count
        const uint32_t c = c_key[i];
        const bool emit = c != LV_EMPTY;
        if (emit) atomicAdd(&hist[hub_mix(c) >> kHubSubBits], 1u);
        wave_append(emit, ((unsigned long long)c << 32) | c_val[i], &s_n, out);
    }
    const uint64_t wsum = group_sum_u64<64>(selfw);
    if ((threadIdx.x & 63) == 0) r_self[threadIdx.x >> 6] = wsum;
    __syncthreads();
    if (run) {
        if (threadIdx.x < kHubParts) {
            const uint32_t h = hist[threadIdx.x];
            if (h)
                a.h.base[((size_t)slot * kHubChunks + q) * kHubParts + threadIdx.x] =
                    atomicAdd(&a.h.cnt[(size_t)slot * kHubParts + threadIdx.x], h);
        }
        if (threadIdx.x == 0) {
            a.h.ccnt[(size_t)slot * kHubChunks + q] = s_n;
            uint64_t t = 0;
            for (int i = 0; i < kBlock / 64; ++i) t += r_self[i];
            if (t) atomicAdd(&a.h.self[slot], (uint32_t)t);
            if (q == 0) {   // post the hub's reduce groups
                const uint32_t groups = hub_groups(a.m.row[v + 1] - a.m.row[v]);
                const uint32_t b0 = atomicAdd(a.h.n_items, groups);
                const uint64_t head = ((uint64_t)v << 32) | ((uint64_t)slot << 8) |
                                      ((uint64_t)ilog2_u32(groups) << kHubPartBits);
                for (uint32_t g = 0; g < groups; ++g) a.h.items[b0 + g] = head | g;
            }
        }
    }
    block_add_stats(a.m.stats, ST_E_ACTIVE, nself, -1, 0, -1, 0);
}

__global__ __launch_bounds__(kBlock) void k_hub_scatter(HubCtx a, const uint32_t *__restrict__ list,
                                                        uint32_t count) {
    __shared__ uint32_t s_off[kHubParts];
    __shared__ uint32_t s_rank[kHubParts];
    const uint32_t q = blockIdx.x % kHubChunks;
    uint32_t v = 0;
    int64_t m2;
    double tw;
    const bool run = hub_runs(a, list, count, blockIdx.x / kHubChunks, v, m2, tw);
    if (!run) return;   // uniform over the workgroup
    const uint32_t slot = a.h.slot_of[v];
    if (threadIdx.x < 64) {
        // start of each partition in the hub's pairs, plus this chunk's run
        const int lane = threadIdx.x;
        uint32_t c = lane < kHubParts ? a.h.cnt[(size_t)slot * kHubParts + lane] : 0u;
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
            if (lane >= off) incl += y;
        }
        if (lane < kHubParts) {
            // base is only written for runs this chunk has; others stay unused
            s_off[lane] = (incl - c) +
                          a.h.base[((size_t)slot * kHubChunks + q) * kHubParts + lane];
            s_rank[lane] = 0;
        }
    }
    __syncthreads();
    const uint64_t beg = a.m.row[v], end = a.m.row[v + 1];
    const uint64_t eo = a.h.ent_off[slot];
    uint64_t cb, ce;
    hub_chunk(beg, end, q, cb, ce);
    const unsigned long long *in = a.h.stage + eo + (cb - beg);
    const uint32_t n = a.h.ccnt[(size_t)slot * kHubChunks + q];
    unsigned long long *pair = a.h.pair + eo;
    for (uint32_t ib = threadIdx.x; ib < n; ib += kBlock * kIlp) {
        unsigned long long pr[kIlp];
#pragma unroll
        for (int u = 0; u < kIlp; ++u) {
            const uint32_t i = ib + (uint32_t)u * kBlock;
            pr[u] = i < n ? __builtin_nontemporal_load(&in[i]) : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kIlp; ++u) {
            if (ib + (uint32_t)u * kBlock >= n) continue;
            const uint32_t p = hub_mix((uint32_t)(pr[u] >> 32)) >> kHubSubBits;
            const uint32_t r = atomicAdd(&s_rank[p], 1u);
            pair[s_off[p] + r] = pr[u];
        }
    }
}

// Persistent over the posted (hub, group) items: a group's pairs are summed
// per community in the LDS table, piece by piece of its partitions' mix range
// (lv_device.h), and every community is scored.
template <int GAIN>
__global__ __launch_bounds__(kBlock) void k_hub_reduce(HubCtx a) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    __shared__ SumTable t;
    __shared__ uint32_t s_beg, s_cnt;
    __shared__ S r_s[kBlock / 64];
    __shared__ uint32_t r_c[kBlock / 64];
    sum_init(t);
    uint64_t ncand = 0;
    const uint32_t n_items = *a.h.n_items;
    for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
        const uint64_t item = a.h.items[it];
        const uint32_t g = (uint32_t)item & (kHubParts - 1);
        const uint32_t v = (uint32_t)(item >> 32);
        const uint32_t slot = ((uint32_t)item >> 8) & 0xFFFFFFu;
        const uint32_t per = (uint32_t)kHubParts >> (((uint32_t)item >> kHubPartBits) & 7u);
        int64_t m2;
        double tw;
        (void)vertex_graph(a.m.bt, v, a.m.m2, a.m.tw, m2, tw);   // items are sweeping hubs only
        lds_barrier();   // the previous item's shared scalars are consumed
        if (threadIdx.x < 64) {
            const uint32_t lane = threadIdx.x;
            const uint32_t c = lane < (uint32_t)kHubParts ? a.h.cnt[(size_t)slot * kHubParts + lane] : 0u;
            const uint64_t before = group_sum_u64<64>(lane < g * per ? c : 0u);
            const uint64_t mine = group_sum_u64<64>(lane >= g * per && lane < (g + 1) * per ? c : 0u);
            if (lane == 0) { s_beg = (uint32_t)before; s_cnt = (uint32_t)mine; }
        }
        lds_barrier();
        const uint32_t own = a.m.comm[v];
        const uint32_t kv = a.m.kdeg[v];
        const uint32_t tot_own = a.m.tot[own];
        const uint32_t self = a.h.self[slot];
        const unsigned long long *pair = a.h.pair + a.h.ent_off[slot] + s_beg;
        const uint32_t np = s_cnt;
        const uint64_t span_lo = (uint64_t)(g * per) << kHubSubBits;
        const uint64_t span_hi = (uint64_t)((g + 1) * per) << kHubSubBits;
        uint64_t width = sum_first_width(span_lo, span_hi, np);
        uint64_t lo = span_lo;
        S best = SC::none();
        uint32_t best_c = a.m.highest ? 0u : LV_EMPTY;
        while (np > 0 && lo < span_hi) {
            const uint64_t rhi = span_hi - lo < width ? span_hi : lo + width;
            const uint32_t nu = sum_piece<kHubIlp>(t, pair, np, lo, rhi);
            if (nu == 0xFFFFFFFFu) {   // too many communities: halve the piece
                width = (rhi - lo) / 2 > 0 ? (rhi - lo) / 2 : 1u;
                continue;
            }
            for (uint32_t ib = threadIdx.x; ib < nu; ib += kBlock * kHubIlp) {
                uint32_t ss[kHubIlp], cs[kHubIlp], ts[kHubIlp];
#pragma unroll
                for (int u = 0; u < kHubIlp; ++u) {
                    const uint32_t i = ib + (uint32_t)u * kBlock;
                    ss[u] = i < nu ? t.used[i] : 0u;
                    cs[u] = i < nu ? t.keys[ss[u]] : LV_EMPTY;
                }
#pragma unroll
                for (int u = 0; u < kHubIlp; ++u) ts[u] = cs[u] != LV_EMPTY ? a.m.tot[cs[u]] : 0u;
#pragma unroll
                for (int u = 0; u < kHubIlp; ++u) {
                    if (cs[u] == LV_EMPTY) continue;
                    ++ncand;
                    const S sc = SC::of(t.vals[ss[u]], self, kv, tot_own, ts[u], m2, tw);
                    if (better_candidate(sc, cs[u], best, best_c, a.m.highest)) { best = sc; best_c = cs[u]; }
                    t.keys[ss[u]] = LV_EMPTY;
                    t.vals[ss[u]] = 0u;
                }
            }
            sum_next_piece(t);
            lo = rhi;
            if (nu < kSumTarget / 2 && width < span_hi - span_lo) width *= 2;   // sparse: widen again
        }
        group_best<64, S>(best, best_c, a.m.highest);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; }
        lds_barrier();
        if (threadIdx.x == 0) {
            for (int i = 0; i < kBlock / 64; ++i)
                if (better_candidate(r_s[i], r_c[i], best, best_c, a.m.highest)) {
                    best = r_s[i];
                    best_c = r_c[i];
                }
            if constexpr (GAIN == 0) {
                a.h.part_s[(size_t)slot * kHubParts + g] = best;
            } else {
                a.h.part_s[(size_t)slot * kHubParts + g] = (int64_t)__float_as_int(best);
            }
            a.h.part_c[(size_t)slot * kHubParts + g] = best_c;
        }
    }
    block_add_stats(a.m.stats, ST_CANDIDATES, ncand, -1, 0, -1, 0);
}

// Fast sweep for hubs: per-part best over entries (no table).
template <int GAIN, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_hub_fast(HubCtx a, const uint32_t *__restrict__ list,
                                                     uint32_t count) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    const uint32_t hi = blockIdx.x / kHubChunks;
    const uint32_t part = blockIdx.x % kHubChunks;
    S best = SC::none();
    uint32_t best_c = LV_EMPTY;
    uint64_t nself = 0;
    uint32_t slot = 0;
    const bool ok = hi < count;
    if (ok) {
        const uint32_t v = list[hi];
        slot = a.h.slot_of[v];
        int64_t m2;
        double tw;
        if (v < a.m.v_end && vertex_graph(a.m.bt, v, a.m.m2, a.m.tw, m2, tw)) {
            const uint32_t kv = a.m.kdeg[v];
            const uint32_t tot_own = a.m.tot[v];
            const uint64_t beg = a.m.row[v], end = a.m.row[v + 1];
            for (uint64_t j = beg + (uint64_t)part * kBlock + threadIdx.x; j < end;
                 j += (uint64_t)kHubChunks * kBlock) {
                const uint32_t x = a.m.col[j];
                if (x == v) continue;
                ++nself;
                const uint32_t wgt = WEIGHTED ? a.m.wt[j] : 1u;
                const S sc = SC::of(wgt, 0u, kv, tot_own, a.m.tot[x], m2, tw);
                if (better_candidate(sc, x, best, best_c, false)) { best = sc; best_c = x; }
            }
        }
    }
    __shared__ S r_s[kBlock / 64];
    __shared__ uint32_t r_c[kBlock / 64];
    group_best<64, S>(best, best_c, false);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; }
    __syncthreads();
    if (threadIdx.x == 0 && ok) {
        for (int i = 0; i < kBlock / 64; ++i)
            if (better_candidate(r_s[i], r_c[i], best, best_c, false)) {
                best = r_s[i];
                best_c = r_c[i];
            }
        if constexpr (GAIN == 0) {
            a.h.part_s[(size_t)slot * kHubParts + part] = best;
        } else {
            a.h.part_s[(size_t)slot * kHubParts + part] = (int64_t)__float_as_int(best);
        }
        a.h.part_c[(size_t)slot * kHubParts + part] = best_c;
        if (part == 0) a.h.self[slot] = 0u;
    }
    block_add_stats(a.m.stats, ST_E_ACTIVE, nself, ST_CANDIDATES, nself, -1, 0);
}

// One wave per hub: reduce the parts, apply the move rule.
template <int GAIN>
__global__ void k_hub_decide(HubCtx a, const uint32_t *__restrict__ list, uint32_t count,
                             int fast) {
    using SC = Score<GAIN>;
    using S = typename SC::T;
    const uint32_t hi = blockIdx.x;
    if (hi >= count) return;
    const uint32_t v = list[hi];
    const uint32_t slot = a.h.slot_of[v];
    const int lane = threadIdx.x;
    const bool highest = fast ? false : a.m.highest;
    const uint32_t parts = fast ? kHubChunks : hub_groups(a.m.row[v + 1] - a.m.row[v]);
    S best = SC::none();
    uint32_t best_c = highest ? 0u : LV_EMPTY;
    if (lane < (int)parts) {
        const int64_t raw = a.h.part_s[(size_t)slot * kHubParts + lane];
        if constexpr (GAIN == 0) best = raw;
        else best = __int_as_float((int)raw);
        best_c = a.h.part_c[(size_t)slot * kHubParts + lane];
    }
    group_best<64, S>(best, best_c, highest);
    int64_t m2;
    double tw;
    if (lane == 0 && v < a.m.v_end && vertex_graph(a.m.bt, v, a.m.m2, a.m.tw, m2, tw)) {
        const uint32_t own = a.m.comm[v];
        const uint32_t kv = a.m.kdeg[v];
        const uint32_t tot_own = a.m.tot[own];
        const uint32_t self = a.h.self[slot];
        const bool has = best != SC::none();
        if (has && SC::positive(best, self, kv, tot_own, m2) && (!fast || v < best_c))
            a.m.target[v] = best_c;
        // a hub row holds > 4096 entries and at most two self-loop entries, so
        // it always has candidates
        atomicAdd(&a.m.stats[ST_CONSIDERED * kStatSlots + (hi & (kStatSlots - 1))], 1ull);
    }
    if (!fast) {   // partition counts, own weight and items start the next sweep at zero
        if (lane < kHubParts) a.h.cnt[(size_t)slot * kHubParts + lane] = 0u;
        if (lane == 0) a.h.self[slot] = 0u;
        if (hi == 0 && lane == 0) *a.h.n_items = 0u;
    }
}

// ------------------------------------------------------------- commit pass

struct CommitCtx {
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *wt;
    const uint32_t *kdeg;
    const uint32_t *comm;
    const uint32_t *target;
    uint32_t *tot;
    uint8_t *mark;
    unsigned long long *stats;
    const uint32_t *gid;          // batches: graph of each vertex (nullptr: one graph)
    unsigned long long *g_dsig;   // batches: sigma_in change per graph
    unsigned long long *ct;       // comm | target << 32 of every vertex (after k_pack_ct);
                                  // kCommitPackedBit marks a vertex in bit 63
};

__global__ void k_pack_ct(const uint32_t *comm, const uint32_t *target, uint32_t n,
                          unsigned long long *ct) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x)
        ct[v] = (unsigned long long)comm[v] | ((unsigned long long)target[v] << 32);
}

// ctot[v] = comm[v] | tot[comm[v]] << 32: the operand of the packed best-move
// tables, built at the start of every hash sweep's best-move pass.
__global__ void k_pack_ctot(const uint32_t *comm, const uint32_t *tot, uint32_t n,
                            unsigned long long *ctot) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = comm[v];
        ctot[v] = (unsigned long long)c | ((unsigned long long)tot[c] << 32);
    }
}

// Commit of mover v (target != comm) over the entries part, part + stride, ...
// of its row in steps of `stride` lanes: tot moves with v's degree (lane 0 of
// part 0), neighbours left in another community are marked, and the change of
// sigma_in from v's entries is returned.
// Neighbour communities: CM == kCommitIdent when the sweep ran at the
// identity partition (every first sweep of a level), so a neighbour's old
// community is the neighbour itself; CM == kCommitPacked reads both
// communities of a neighbour from one 8-byte word (ct = comm | target << 32).
// CM == kCommitPackedBit also keeps the neighbour's mark in bit 63 of its
// word (bit 31 of the target, free while ids are below 2^31): the mark costs
// no access of its own, and is set by storing the word's high half with the
// bit, the same value from every writer.  k_finalize reads it back.
constexpr int kCommitIdent = 1;
constexpr int kCommitPacked = 2;
constexpr int kCommitPackedBit = 3;
constexpr uint32_t kMarkBitMaxN = 0x80000000u;   // the bit path needs every id < 2^31

// The caller loads v's words (communities, row bounds, degree) together, so
// their latencies overlap.
template <bool WEIGHTED, int CM, int ILP = kIlp>
__device__ inline int64_t commit_row(const CommitCtx &a, uint32_t v, uint32_t own, uint32_t c,
                                     uint64_t beg, uint64_t end, uint32_t kv, uint32_t first,
                                     uint32_t stride) {
    if (first == 0) {
        atomicAdd(&a.tot[c], kv);
        atomicSub(&a.tot[own], kv);
    }
    int64_t dsig = 0;
    for (uint64_t jb = beg + first; jb < end; jb += (uint64_t)stride * ILP) {
        uint32_t xs[ILP], ws[ILP], ox[ILP], nx[ILP];
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            const uint64_t j = jb + (uint64_t)u * stride;
            xs[u] = j < end ? stream_load(&a.col[j]) : v;
            ws[u] = (WEIGHTED && j < end) ? stream_load(&a.wt[j]) : 1u;
        }
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            if (CM == kCommitIdent) {
                ox[u] = xs[u];
                nx[u] = xs[u] != v ? a.target[xs[u]] : 0u;
            } else {
                const unsigned long long p = xs[u] != v ? a.ct[xs[u]] : 0ull;
                ox[u] = (uint32_t)p;
                nx[u] = (uint32_t)(p >> 32);   // CM == kCommitPackedBit: with the mark
            }
        }
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            if (xs[u] == v) continue;
            const uint32_t hi = nx[u];
            if (CM == kCommitPackedBit) nx[u] &= 0x7FFFFFFFu;
            const int d = (int)(nx[u] == c) - (int)(ox[u] == own);
            if (d) dsig += (int64_t)ws[u] * d * (ox[u] != nx[u] ? 1 : 2);
            // most neighbours are marked by several movers: store only the
            // first mark, so lines are not dirtied and written back again and
            // again
            if (nx[u] == c) continue;
            if (CM == kCommitPackedBit) {
                if (!(hi >> 31)) reinterpret_cast<uint32_t *>(&a.ct[xs[u]])[1] = hi | 0x80000000u;
            } else if (!a.mark[xs[u]]) {
                a.mark[xs[u]] = 1;
            }
        }
    }
    return dsig;
}

// A group of G (<= 64) lanes per listed vertex of buckets 0-6, all seven
// buckets in one launch: workgroup b takes bucket q with wg[q] <= b < wg[q+1].
constexpr int kSmallBuckets = kNumBuckets - kBigBuckets;
// Lanes per vertex: a few entries per lane (up to 4-16, one to four kIlp
// steps), so a wave commits many vertices and keeps many independent chains
// of row, neighbour word and mark in flight (measured: {4, 8, 16, 32, 64, 64},
// one lane per entry, was 15% slower on the 4,096-graph batch).  Bucket 6
// (257-1024 entries, 64 lanes) is committed here rather than as compacted
// movers of a whole workgroup each (measured slightly faster).
constexpr int kSmallG[kSmallBuckets] = {2, 2, 4, 8, 8, 16, 64};

struct SmallLists {
    const uint32_t *list[kSmallBuckets];
    uint32_t count[kSmallBuckets];
    uint32_t wg[kSmallBuckets + 1];   // first workgroup of each bucket
};

template <int G, bool WEIGHTED, int CM>
__device__ inline int64_t commit_group(const CommitCtx &a, const uint32_t *__restrict__ list,
                                       uint32_t count, uint32_t wb, uint32_t &v,
                                       uint64_t &ment) {
    constexpr int GPB = kBlock / G;
    const int lane = threadIdx.x & (G - 1);
    const uint32_t idx = wb * GPB + group_index<G>();
    int64_t dsig = 0;
    if (idx < count) {
        v = list[idx];
        const uint32_t own = a.comm[v];
        const uint32_t c = a.target[v];
        const uint64_t beg = a.row[v], end = a.row[v + 1];
        const uint32_t kv = a.kdeg[v];
        if (c != own) {
            // rows of up to 16 entries per lane (groups of 16 and 64 lanes):
            // eight entries per lane in flight, half the round trips of four
            dsig = commit_row<WEIGHTED, CM, (G >= 16 ? 2 * kIlp : kIlp)>(a, v, own, c, beg, end, kv,
                                                                       (uint32_t)lane, (uint32_t)G);
            if (lane == 0) ment = end - beg;
        }
    }
    return dsig;
}

template <bool WEIGHTED, int CM>
__global__ __launch_bounds__(kBlock) void k_commit(CommitCtx a, SmallLists sl) {
    uint32_t q = 0;
    while (q + 1 < (uint32_t)kSmallBuckets && blockIdx.x >= sl.wg[q + 1]) ++q;
    const uint32_t wb = blockIdx.x - sl.wg[q];
    uint32_t v = LV_EMPTY;
    int64_t dsig = 0;
    uint64_t ment = 0;
    switch (kSmallG[q]) {   // uniform over the workgroup
        case 2: dsig = commit_group<2, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
        case 4: dsig = commit_group<4, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
        case 8: dsig = commit_group<8, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
        case 16: dsig = commit_group<16, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
        case 32: dsig = commit_group<32, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
        default: dsig = commit_group<64, WEIGHTED, CM>(a, sl.list[q], sl.count[q], wb, v, ment); break;
    }
    if (a.gid) wave_keyed_add(a.g_dsig, v != LV_EMPTY ? a.gid[v] : 0u, (uint64_t)dsig, dsig != 0);
    block_add_stats(a.stats, ST_DSIGMA, (uint64_t)dsig, ST_MOVER_ENTRIES, ment, -1, 0);
}

// Large rows (bucket 7, hubs): most listed vertices stay after the
// first sweep, so the movers are compacted first (per bucket, one atomic per
// wave) and the commit runs over (mover, part) items on a persistent grid.
// Parts per mover (kBlock lanes each): a hub mover of ~9,000 entries then
// gives its lanes four or five entries each (measured: 8 parts committed
// R-MAT-22 slightly faster than 16 or 32).
constexpr uint32_t kBigParts[kBigBuckets] = {4u, 8u};
// Grid of k_commit_big: many rounds of resident workgroups, for the same
// reason (measured faster than 2,048 to 16,384).
constexpr uint32_t kCommitBlocks = 32768;

struct BigLists {
    const uint32_t *list[kBigBuckets];
    uint32_t count[kBigBuckets];
    uint32_t off[kBigBuckets];       // region of each bucket's movers in `movers`
};

__global__ __launch_bounds__(kBlock) void k_big_movers(BigLists bl, const uint32_t *comm,
                                                       const uint32_t *target, uint32_t *movers,
                                                       uint32_t *n_movers) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t b = 0, k = i;
    while (b < kBigBuckets && k >= bl.count[b]) { k -= bl.count[b]; ++b; }
    uint32_t v = 0;
    bool mv = false;
    if (b < kBigBuckets) {
        v = bl.list[b][k];
        mv = target[v] != comm[v];
    }
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)kBigBuckets; ++q) {
        const bool sel = mv && b == q;
        const unsigned long long m = __ballot(sel);
        if (m == 0) continue;
        const int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(&n_movers[q], (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, leader);
        if (sel) movers[bl.off[q] + base + (uint32_t)__popcll(m & lanemask_lt())] = v;
    }
}

template <bool WEIGHTED, int CM>
__global__ __launch_bounds__(kBlock) void k_commit_big(CommitCtx a, BigLists bl,
                                                       const uint32_t *__restrict__ movers,
                                                       const uint32_t *__restrict__ n_movers) {
    uint32_t nm[kBigBuckets];
    uint32_t total = 0;
#pragma unroll
    for (int q = 0; q < kBigBuckets; ++q) {
        nm[q] = n_movers[q];
        total += nm[q] * kBigParts[q];
    }
    int64_t dsum = 0;
    uint64_t ment = 0;
    for (uint32_t it = blockIdx.x; it < total; it += gridDim.x) {
        uint32_t q = 0, j = it;
        while (j >= nm[q] * kBigParts[q]) { j -= nm[q] * kBigParts[q]; ++q; }
        const uint32_t v = movers[bl.off[q] + j / kBigParts[q]];
        const uint32_t part = j % kBigParts[q];
        const int64_t dsig = commit_row<WEIGHTED, CM>(a, v, a.comm[v], a.target[v], a.row[v],
                                                  a.row[v + 1], a.kdeg[v],
                                                  part * kBlock + threadIdx.x,
                                                  kBlock * kBigParts[q]);
        if (a.gid) wave_keyed_add(a.g_dsig, a.gid[v], (uint64_t)dsig, dsig != 0);
        dsum += dsig;
        if (part == 0 && threadIdx.x == 0) ment += a.row[v + 1] - a.row[v];
    }
    block_add_stats(a.stats, ST_DSIGMA, (uint64_t)dsum, ST_MOVER_ENTRIES, ment, -1, 0);
}

// ----------------------------------------------------------- finalize pass

// Each workgroup takes kFinTiles consecutive tiles of kBlock vertices and
// reserves its run of every next list with one atomic per bucket.
constexpr int kFinTiles = 16;

__global__ __launch_bounds__(kBlock) void k_finalize(uint32_t *comm, const uint32_t *target,
                                                     const uint32_t *tot, uint8_t *mark,
                                                     const uint8_t *bucket, uint32_t n,
                                                     ListPtrs next, uint32_t *next_counts,
                                                     unsigned long long *stats,
                                                     const uint32_t *gid,
                                                     unsigned long long *g_sumsq,
                                                     const unsigned long long *ctm) {
    __shared__ uint32_t s_cnt[kNumBuckets];
    __shared__ uint32_t s_base[kNumBuckets];
    if (threadIdx.x < kNumBuckets) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    uint64_t moved = 0, sq = 0;
    const uint32_t v0 = blockIdx.x * (uint32_t)(kBlock * kFinTiles);
    uint32_t rank[kFinTiles];
    uint8_t bk[kFinTiles];
#pragma unroll
    for (int i = 0; i < kFinTiles; ++i) {
        const uint32_t v = v0 + (uint32_t)i * kBlock + threadIdx.x;
        rank[i] = LV_EMPTY;
        bk[i] = 0;
        uint64_t q = 0;
        uint32_t g = 0;
        if (v < n) {
            const uint32_t tv = target[v];
            if (tv != comm[v]) { comm[v] = tv; ++moved; }
            q = square_u64(tot[v]);
            sq += q;
            if (gid) g = gid[v];
            // marks of a kCommitPackedBit commit are in bit 63 of ctm
            const bool in_ct = ctm && (ctm[v] >> 63);
            const uint8_t mb = mark[v];
            if (mb || in_ct) {
                if (mb) mark[v] = 0;
                bk[i] = bucket[v];
                rank[i] = atomicAdd(&s_cnt[bk[i]], 1u);
            }
        }
        if (gid) wave_keyed_add(g_sumsq, g, q, q != 0);   // gid is uniform: whole waves call it
    }
    __syncthreads();
    if (threadIdx.x < kNumBuckets) {
        const uint32_t c = s_cnt[threadIdx.x];
        s_base[threadIdx.x] = c ? atomicAdd(&next_counts[threadIdx.x], c) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kFinTiles; ++i)
        if (rank[i] != LV_EMPTY)
            next.p[bk[i]][s_base[bk[i]] + rank[i]] = v0 + (uint32_t)i * kBlock + threadIdx.x;
    block_add_stats(stats, ST_MOVED, moved, ST_SUMSQ, sq, -1, 0);
}

// ------------------------------------------------------------- aggregation

__global__ void k_renumber_flags(const uint32_t *tot, uint32_t n, uint32_t *flags) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x)
        flags[c] = count_not_zero(tot[c]);
}

__global__ void k_update_best(uint32_t *best, uint32_t n0, const uint32_t *comm,
                              const uint32_t *map) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n0;
         v += (size_t)gridDim.x * blockDim.x)
        best[v] = renumbered(map, comm[best[v]]);
}

__global__ void k_new_degrees(const uint32_t *tot, const uint32_t *map, uint32_t n,
                              uint32_t *k_new) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x)
        if (tot[c] > 0) k_new[renumbered(map, (uint32_t)c)] = tot[c];
}

// ------------------------------------------------------- batches of graphs

__global__ __launch_bounds__(kBlock) void k_batch_comm_counts(const uint32_t *tot,
                                                              const uint32_t *gid, uint32_t n,
                                                              unsigned long long *counts) {
    const uint32_t ntiles = (n + kBlock - 1) / kBlock;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t c = t * kBlock + threadIdx.x;
        const bool has = c < n && tot[c] > 0;
        wave_keyed_add(counts, has ? gid[c] : 0u, 1ull, has);
    }
}

__global__ void k_batch_next_gid(const uint32_t *tot, const uint32_t *map, const uint32_t *gid,
                                 uint32_t n, uint32_t *gid_new) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
         c += (size_t)gridDim.x * blockDim.x)
        if (tot[c] > 0) gid_new[renumbered(map, (uint32_t)c)] = gid[c];
}

}  // namespace

// ====================================================================== host

void batch_comm_counts(const uint32_t *tot, const uint32_t *gid, uint32_t n,
                       unsigned long long *counts, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_batch_comm_counts, dim3(grid_for(n, kBlock * 4, 16384u)), dim3(kBlock), 0,
                       s, tot, gid, n, counts);
}

void batch_next_gid(const uint32_t *tot, const uint32_t *map, const uint32_t *gid, uint32_t n,
                    uint32_t *gid_new, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_batch_next_gid, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, tot,
                       map, gid, n, gid_new);
}

void sweep_reset(unsigned long long *stats, size_t n_stats, uint32_t *n_movers,
                 uint32_t *next_counts, hipStream_t s) {
    static_assert(kBigBuckets <= kBlock && kNumBuckets <= kBlock, "one counter per thread");
    hipLaunchKernelGGL(k_sweep_reset, dim3(grid_for(n_stats, kBlock * 4)), dim3(kBlock), 0, s,
                       stats, n_stats, n_movers, (uint32_t)kBigBuckets, next_counts,
                       (uint32_t)kNumBuckets);
}

void fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, v, n);
}
void fill_u8(uint8_t *p, uint8_t v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, kBlock * 16)), dim3(kBlock), 0, s, p, v, n);
}
void fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, v, n);
}
void iota_u32(uint32_t *p, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, p, n);
}

void scatter_u32(uint32_t *dst, const uint32_t *idx, const uint32_t *val, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_scatter_u32, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, dst, idx,
                       val, n);
}

void stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / sizeof(uint4);
    if (!n) return;
    hipLaunchKernelGGL(k_stream_copy, dim3(grid_for(n, kBlock * 8, 1u << 20)), dim3(kBlock), 0, s,
                       (const uint4 *)src, (uint4 *)dst, n);
}

void gather_probe(const uint32_t *col, uint64_t E, const void *arr, int width,
                  unsigned long long *out, hipStream_t s) {
    static_assert(kGatherProbeWords == (size_t)kGatherProbeBlocks * kBlock, "a word per lane");
    if (!E) return;
    if (width == 8)
        hipLaunchKernelGGL(k_gather_probe<unsigned long long>, dim3(kGatherProbeBlocks),
                           dim3(kBlock), 0, s, col, E, (const unsigned long long *)arr, out);
    else
        hipLaunchKernelGGL(k_gather_probe<uint32_t>, dim3(kGatherProbeBlocks), dim3(kBlock), 0, s,
                           col, E, (const uint32_t *)arr, out);
}

void nonself_counts(const uint64_t *row, const uint32_t *col, uint32_t n, uint32_t *nsel,
                    hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_nonself_counts, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, row, col, n,
                       nsel);
}

void chunk_selected(const uint32_t *list, uint32_t count, const uint32_t *nsel, const uint32_t *cv,
                    uint32_t nc, unsigned long long *csel, hipStream_t s) {
    if (!count) return;
    hipLaunchKernelGGL(k_chunk_selected, dim3(grid_for(count, kBlock * 8, 4096u)), dim3(kBlock), 0,
                       s, list, count, nsel, cv, nc, csel);
}

void coo_sources(const uint64_t *row, uint32_t n, uint32_t *src, uint32_t *deg, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_coo_sources, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, row, n, src,
                       deg);
}

void gather_degrees(const uint64_t *row, const uint32_t *ids, uint32_t n, uint64_t *deg,
                    hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_gather_degrees, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, row, ids, n,
                       deg);
}

void community_totals(const uint32_t *comm, const uint32_t *kdeg, uint32_t n, uint32_t *tot,
                      hipStream_t s) {
    fill_u32(tot, 0u, n, s);
    if (!n) return;
    hipLaunchKernelGGL(k_community_totals, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, comm,
                       kdeg, n, tot);
}

size_t scan_scratch_elems(size_t n) {
    size_t total = 0;
    size_t m = n;
    while (m > (size_t)kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        total += m;
    }
    return total + 1;
}

void scan_u32(const uint32_t *in, uint32_t *out, size_t n, bool inclusive, uint32_t *scratch,
              hipStream_t s) {
    scan_impl<uint32_t>(in, out, n, inclusive, scratch, s);
}
void scan_u64(const uint64_t *in, uint64_t *out, size_t n, bool inclusive, uint64_t *scratch,
              hipStream_t s) {
    scan_impl<uint64_t>(in, out, n, inclusive, scratch, s);
}

size_t radix_hist_elems(size_t n) {
    const size_t nb = (n + kRsTile - 1) / kRsTile;
    return 256 * (nb ? nb : 1);
}

void radix_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, size_t n, int bits,
                      uint32_t *hist, uint32_t *scan_scratch, hipStream_t s) {
    if (n < 2 || bits <= 0) return;
    const uint32_t nb = (uint32_t)((n + kRsTile - 1) / kRsTile);
    const size_t hn = (size_t)256 * nb;
    uint64_t *ki = k0, *ko = k1;
    uint32_t *vi = v0, *vo = v1;
    int passes = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_upsweep, dim3(nb), dim3(kBlock), 0, s, (const uint64_t *)ki, n,
                           shift, hist, nb);
        scan_u32(hist, hist, hn, false, scan_scratch, s);
        hipLaunchKernelGGL(k_rs_downsweep, dim3(nb), dim3(kBlock), 0, s, (const uint64_t *)ki,
                           (const uint32_t *)vi, ko, vo, n, shift, (const uint32_t *)hist, nb);
        uint64_t *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
        ++passes;
    }
    if (passes & 1) {
        (void)hipMemcpyAsync(k0, ki, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(v0, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
}

bool bounds_valid(const uint32_t *max_deg) {
    if (max_deg[kNumBuckets - 2] < 2) return false;
    for (int b = 0; b < kNumBuckets - 1; ++b) {
        if (max_deg[b] > (b == 5 ? kWide5MaxDeg : kBucketMaxDeg[b])) return false;
        if (b > 0 && max_deg[b] < max_deg[b - 1]) return false;
    }
    return true;
}

void bucket_ids(const uint64_t *row, uint32_t n, uint8_t *bucket, const BucketBounds &bounds,
                hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_bucket_ids, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, row, n,
                       bucket, bounds);
}

void bucket_counts(const uint8_t *bucket, uint32_t n, uint32_t *counts, hipStream_t s) {
    fill_u32(counts, 0u, kNumBuckets, s);
    if (!n) return;
    hipLaunchKernelGGL(k_bucket_counts, dim3(grid_for(n, kBlock * 16, 4096u)), dim3(kBlock), 0, s,
                       bucket, n, counts);
}

void build_lists_all(const uint8_t *bucket, uint32_t n, uint32_t *const *lists, uint32_t *counts,
                     hipStream_t s) {
    ListPtrs lp;
    for (int b = 0; b < kNumBuckets; ++b) lp.p[b] = lists[b];
    fill_u32(counts, 0u, kNumBuckets, s);
    if (!n) return;
    const uint32_t per = (uint32_t)(kBlock * kListTiles);
    hipLaunchKernelGGL(k_build_lists_all, dim3((n + per - 1) / per), dim3(kBlock), 0, s, bucket, n,
                       lp, counts);
}

void identity_parts(const LevelView &g, unsigned long long *stats, hipStream_t s) {
    if (!g.n) return;
    hipLaunchKernelGGL(k_identity_parts, dim3(grid_for(g.n, kBlock, 8192u)), dim3(kBlock), 0, s, g,
                       stats);
}

void modularity_parts(const LevelView &g, const uint32_t *comm, const uint32_t *tot,
                      unsigned long long *stats, hipStream_t s) {
    const uint64_t work = (g.E + kChunk - 1) / kChunk > g.n ? (g.E + kChunk - 1) / kChunk : g.n;
    hipLaunchKernelGGL(k_modularity_parts, dim3(grid_for(work, kBlock, 8192u)), dim3(kBlock), 0, s,
                       g, comm, tot, stats);
}

namespace {

// Tables of up to kPackedMaxT slots are packed; bucket 7's keeps the
// per-candidate total gathers rather than 50% more LDS per vertex.
// Workgroups of `kernel` the device holds at once (CUs x occupancy), cached
// in `slot` at the first launch (every device in use is the same model).
// The persistent kernels launch kPersistRounds times that: each workgroup
// still walks several vertices with the next one in flight, and the uneven
// rows are balanced in rounds (measured: 8 rounds faster than 1, 4 and 16).
constexpr uint32_t kPersistRounds = 8;
template <typename K>
uint32_t resident_blocks(K kernel, std::atomic<uint32_t> &slot) {
    uint32_t r = slot.load(std::memory_order_relaxed);
    if (r) return r;
    int dev = 0, cus = 0, per = 0;
    LV_HIP(hipGetDevice(&dev));
    LV_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, 0));
    r = (uint32_t)std::max(1, cus * std::max(per, 1));
    slot.store(r, std::memory_order_relaxed);
    return r;
}

template <int G, int T, int GAIN, bool WEIGHTED>
void launch_move_table_k(const MoveCtx &m, const uint32_t *list, uint32_t count, hipStream_t s) {
    constexpr int GPB = kBlock / G;
    constexpr bool PK = T <= kPackedMaxT;
    auto kernel = k_move_table<G, T, GAIN, WEIGHTED, PK>;
    uint32_t blocks = (count + GPB - 1) / GPB;
    if (G == kBlock) {   // persistent: a few rounds of resident workgroups
        static std::atomic<uint32_t> resident{0};
        blocks = std::min(blocks, kPersistRounds * resident_blocks(kernel, resident));
    }
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, s, m, list, count);
}

template <int G, int T>
void launch_move_table(const MoveCtx &m, bool weighted, int gain, const uint32_t *list,
                       uint32_t count, hipStream_t s) {
    if (gain == 0) {
        if (weighted) launch_move_table_k<G, T, 0, true>(m, list, count, s);
        else launch_move_table_k<G, T, 0, false>(m, list, count, s);
    } else {
        if (weighted) launch_move_table_k<G, T, 1, true>(m, list, count, s);
        else launch_move_table_k<G, T, 1, false>(m, list, count, s);
    }
}

template <int G>
void launch_move_fast(const MoveCtx &m, bool weighted, int gain, const uint32_t *list,
                      uint32_t count, hipStream_t s) {
    constexpr int GPB = kBlock / G;
    const dim3 grid((count + GPB - 1) / GPB), blk(kBlock);
    if (gain == 0) {
        if (weighted) hipLaunchKernelGGL((k_move_fast<G, 0, true>), grid, blk, 0, s, m, list, count);
        else hipLaunchKernelGGL((k_move_fast<G, 0, false>), grid, blk, 0, s, m, list, count);
    } else {
        if (weighted) hipLaunchKernelGGL((k_move_fast<G, 1, true>), grid, blk, 0, s, m, list, count);
        else hipLaunchKernelGGL((k_move_fast<G, 1, false>), grid, blk, 0, s, m, list, count);
    }
}

// The (G, T) instantiations behind kTileG.
constexpr int kTableT5Wide = (int)table_slots(kWide5MaxDeg);
static_assert(kTableT5Wide == 772 && kTableT5Wide <= kPackedMaxT, "wide bucket-5 table");

void launch_table_tile(int b, int G, const MoveCtx &m, bool wtd, int gain, const uint32_t *L,
                       uint32_t cnt, hipStream_t s, bool wide5) {
#define LV_T(GG, TT) launch_move_table<GG, TT>(m, wtd, gain, L, cnt, s)
    switch (b) {
        case 0: if (G == 1) LV_T(1, kTableT[0]); else if (G == 4) LV_T(4, kTableT[0]); else LV_T(2, kTableT[0]); break;
        case 1: if (G == 2) LV_T(2, kTableT[1]); else if (G == 8) LV_T(8, kTableT[1]); else LV_T(4, kTableT[1]); break;
        case 2: if (G == 4) LV_T(4, kTableT[2]); else if (G == 16) LV_T(16, kTableT[2]); else LV_T(8, kTableT[2]); break;
        case 3: if (G == 8) LV_T(8, kTableT[3]); else if (G == 32) LV_T(32, kTableT[3]); else LV_T(16, kTableT[3]); break;
        case 4: if (G == 16) LV_T(16, kTableT[4]); else if (G == 64) LV_T(64, kTableT[4]); else LV_T(32, kTableT[4]); break;
        case 5:
            if (wide5) {
                if (G == 32) LV_T(32, kTableT5Wide); else if (G == 256) LV_T(256, kTableT5Wide); else LV_T(64, kTableT5Wide);
            } else {
                if (G == 32) LV_T(32, kTableT[5]); else if (G == 256) LV_T(256, kTableT[5]); else LV_T(64, kTableT[5]);
            }
            break;
        case 6: LV_T(256, kTableT[6]); break;
        default: LV_T(256, kMoveTableT[7]); break;
    }
#undef LV_T
}

void launch_fast_tile(int G, const MoveCtx &m, bool wtd, int gain, const uint32_t *L, uint32_t cnt,
                      hipStream_t s) {
    switch (G) {
        case 1: launch_move_fast<1>(m, wtd, gain, L, cnt, s); break;
        case 2: launch_move_fast<2>(m, wtd, gain, L, cnt, s); break;
        case 4: launch_move_fast<4>(m, wtd, gain, L, cnt, s); break;
        case 8: launch_move_fast<8>(m, wtd, gain, L, cnt, s); break;
        case 16: launch_move_fast<16>(m, wtd, gain, L, cnt, s); break;
        case 32: launch_move_fast<32>(m, wtd, gain, L, cnt, s); break;
        case 64: launch_move_fast<64>(m, wtd, gain, L, cnt, s); break;
        default: launch_move_fast<256>(m, wtd, gain, L, cnt, s); break;
    }
}

MoveCtx move_ctx(const SweepArgs &a) {
    MoveCtx m;
    m.row = a.g.row; m.col = a.g.col; m.wt = a.g.wt; m.kdeg = a.g.kdeg;
    m.comm = a.comm; m.tot = a.tot; m.target = a.target;
    m.ctot = a.ct;
    m.m2 = a.m2; m.tw = a.tw; m.v_end = a.v_end; m.highest = a.highest;
    m.bt = a.bt;
    return m;
}

}  // namespace

void pack_comm_tot(const SweepArgs &a, hipStream_t s) {
    if (!a.g.n) return;
    hipLaunchKernelGGL(k_pack_ctot, dim3(grid_for(a.g.n, kBlock * 4)), dim3(kBlock), 0, s, a.comm,
                       a.tot, a.g.n, a.ct);
}

void move_bucket(const SweepArgs &a, int b, int sel, const uint32_t *list, uint32_t count,
                 hipStream_t s) {
    if (!count || b >= kNumBuckets - 1) return;
    MoveCtx m = move_ctx(a);
    m.stats = a.bstats + (size_t)b * kNumStats * kStatSlots;
    const int G = kTileG[b][sel];
    if (a.kind == 0) launch_fast_tile(G, m, a.g.wt != nullptr, a.gain, list, count, s);
    else launch_table_tile(b, G, m, a.g.wt != nullptr, a.gain, list, count, s, a.wide5);
}

void move_pass(const SweepArgs &a, const HubTables &hub, uint32_t *const *lists,
               const uint32_t *counts, hipStream_t s0, const SideStreams *side) {
    MoveCtx m = move_ctx(a);
    const bool wtd = a.g.wt != nullptr;
    const bool fast = a.kind == 0;
    // Buckets 5, 7 and the hubs run on side streams beside buckets 0-4 (their
    // tails and launch gaps overlap); bucket 6, the largest launch, runs after
    // the join with the device to itself.  A bucket's profiling events go on
    // the stream it runs on.
    if (!fast) pack_comm_tot(a, s0);   // before the fork: every stream reads it
    const bool fork = side && side->ready();
    if (fork) {
        LV_HIP(hipEventRecord(side->fork, s0));
        for (int i = 0; i < 3; ++i) LV_HIP(hipStreamWaitEvent(side->s[i], side->fork, 0));
    }
    static constexpr int kOrder[kNumBuckets] = {0, 1, 2, 3, 4, 5, 7, kHubBucket, 6};
    for (int k = 0; k < kNumBuckets; ++k) {
        const int b = kOrder[k];
        if (fork && b == 6)
            for (int i = 0; i < 3; ++i) {
                LV_HIP(hipEventRecord(side->join[i], side->s[i]));
                LV_HIP(hipStreamWaitEvent(s0, side->join[i], 0));
            }
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        const uint32_t *L = lists[b];
        const hipStream_t s = !fork ? s0
                              : b == 5 ? side->s[2]
                              : b == 7 ? side->s[0]
                              : b == kHubBucket ? side->s[1]
                              : s0;
        m.stats = a.bstats + (size_t)b * kNumStats * kStatSlots;
        const bool timed = a.bev && (a.bev_mask >> b & 1u);
        if (timed) (void)hipEventRecord(a.bev[2 * b], s);
        if (fast) {
            switch (b) {
                case 0: case 1: case 2: case 3: case 4: case 5: case 6: case 7:
                    launch_fast_tile(kTileG[b][a.tiles.sel[b][1]], m, wtd, a.gain, L, cnt, s);
                    break;
                default: {
                    HubCtx h{m, hub};
                    const dim3 grid(cnt * kHubChunks), blk(kBlock);
                    if (a.gain == 0) {
                        if (wtd) hipLaunchKernelGGL((k_hub_fast<0, true>), grid, blk, 0, s, h, L, cnt);
                        else hipLaunchKernelGGL((k_hub_fast<0, false>), grid, blk, 0, s, h, L, cnt);
                        hipLaunchKernelGGL(k_hub_decide<0>, dim3(cnt), dim3(64), 0, s, h, L, cnt, 1);
                    } else {
                        if (wtd) hipLaunchKernelGGL((k_hub_fast<1, true>), grid, blk, 0, s, h, L, cnt);
                        else hipLaunchKernelGGL((k_hub_fast<1, false>), grid, blk, 0, s, h, L, cnt);
                        hipLaunchKernelGGL(k_hub_decide<1>, dim3(cnt), dim3(64), 0, s, h, L, cnt, 1);
                    }
                    // stats: considered hubs counted host-side from the list size
                    break;
                }
            }
        } else {
            switch (b) {
                case 0: case 1: case 2: case 3: case 4: case 5: case 6: case 7:
                    launch_table_tile(b, kTileG[b][a.tiles.sel[b][0]], m, wtd, a.gain, L, cnt, s,
                                      a.wide5);
                    break;
                default: {
                    HubCtx h{m, hub};
                    const dim3 grid(cnt * kHubChunks), blk(kBlock);
                    if (wtd) {
                        hipLaunchKernelGGL(k_hub_count<true>, grid, blk, 0, s, h, L, cnt);
                        hipLaunchKernelGGL(k_hub_scatter, grid, blk, 0, s, h, L, cnt);
                    } else {
                        hipLaunchKernelGGL(k_hub_count<false>, grid, blk, 0, s, h, L, cnt);
                        hipLaunchKernelGGL(k_hub_scatter, grid, blk, 0, s, h, L, cnt);
                    }
                    // persistent: the items posted by k_hub_count
                    const dim3 rgrid(std::min<uint32_t>(cnt * kHubParts, kHubReduceBlocks));
                    if (a.gain == 0) {
                        hipLaunchKernelGGL(k_hub_reduce<0>, rgrid, blk, 0, s, h);
                        hipLaunchKernelGGL(k_hub_decide<0>, dim3(cnt), dim3(64), 0, s, h, L, cnt, 0);
                    } else {
                        hipLaunchKernelGGL(k_hub_reduce<1>, rgrid, blk, 0, s, h);
                        hipLaunchKernelGGL(k_hub_decide<1>, dim3(cnt), dim3(64), 0, s, h, L, cnt, 0);
                    }
                    break;
                }
            }
        }
        if (timed) (void)hipEventRecord(a.bev[2 * b + 1], s);
    }
}

static void commit_big_rows(const SweepArgs &a, const CommitCtx &c, uint32_t *const *lists,
                            const uint32_t *counts, bool wtd, bool ident, bool bit, hipStream_t s);

static std::atomic<bool> g_mark_bytes{false};
void set_mark_bytes(bool on) { g_mark_bytes.store(on); }

// While every id is below 2^31 every commit is packed and keeps its marks in
// bit 63 of ct (at the identity partition too: a neighbour's word then holds
// the neighbour itself and its target).
static bool marks_in_ct(const SweepArgs &a) {
    return a.ct != nullptr && a.g.n <= kMarkBitMaxN && !g_mark_bytes.load();
}

void commit_pass(const SweepArgs &a, uint32_t *const *lists, const uint32_t *counts,
                 hipStream_t s, const SideStreams *side) {
    CommitCtx c;
    c.row = a.g.row; c.col = a.g.col; c.wt = a.g.wt; c.kdeg = a.g.kdeg;
    c.comm = a.comm; c.target = a.target; c.tot = a.tot; c.mark = a.mark; c.stats = a.stats;
    c.gid = a.bt.gid; c.g_dsig = a.bt.dsig;
    const bool wtd = a.g.wt != nullptr;
    const bool bit = marks_in_ct(a);
    // fast sweeps run at the identity partition: without the mark bit their
    // commit reads a neighbour's target only
    const bool ident = a.kind == 0 && !bit;
    const dim3 blk(kBlock);
    c.ct = a.ct;
    if (!ident && a.g.n)
        hipLaunchKernelGGL(k_pack_ct, dim3(grid_for(a.g.n, kBlock * 4)), blk, 0, s, a.comm,
                           (const uint32_t *)a.target, a.g.n, a.ct);
    SmallLists sl;
    uint32_t wgs = 0;
    for (int q = 0; q < kSmallBuckets; ++q) {
        sl.list[q] = lists[q];
        sl.count[q] = counts[q];
        sl.wg[q] = wgs;
        const uint32_t gpb = kBlock / kSmallG[q];
        wgs += (counts[q] + gpb - 1) / gpb;
    }
    sl.wg[kSmallBuckets] = wgs;
    // buckets 0-6 on a side stream beside the large rows (both after k_pack_ct;
    // their writes are atomics or the same mark bytes)
    const bool fork = wgs && side && side->ready();
    const hipStream_t ss = fork ? side->s[0] : s;
    if (fork) {
        LV_HIP(hipEventRecord(side->fork, s));
        LV_HIP(hipStreamWaitEvent(ss, side->fork, 0));
    }
    if (wgs) {
        if (ident) {
            if (wtd) hipLaunchKernelGGL((k_commit<true, kCommitIdent>), dim3(wgs), blk, 0, ss, c, sl);
            else hipLaunchKernelGGL((k_commit<false, kCommitIdent>), dim3(wgs), blk, 0, ss, c, sl);
        } else if (bit) {
            if (wtd) hipLaunchKernelGGL((k_commit<true, kCommitPackedBit>), dim3(wgs), blk, 0, ss, c, sl);
            else hipLaunchKernelGGL((k_commit<false, kCommitPackedBit>), dim3(wgs), blk, 0, ss, c, sl);
        } else {
            if (wtd) hipLaunchKernelGGL((k_commit<true, kCommitPacked>), dim3(wgs), blk, 0, ss, c, sl);
            else hipLaunchKernelGGL((k_commit<false, kCommitPacked>), dim3(wgs), blk, 0, ss, c, sl);
        }
    }
    commit_big_rows(a, c, lists, counts, wtd, ident, bit, s);
    if (fork) {
        LV_HIP(hipEventRecord(side->join[0], ss));
        LV_HIP(hipStreamWaitEvent(s, side->join[0], 0));
    }
}

// Large rows (bucket 7, hubs): movers compacted, then the persistent commit.
static void commit_big_rows(const SweepArgs &a, const CommitCtx &c, uint32_t *const *lists,
                            const uint32_t *counts, bool wtd, bool ident, bool bit, hipStream_t s) {
    const dim3 blk(kBlock);
    BigLists bl;
    uint32_t listed = 0, items = 0;
    for (int q = 0; q < kBigBuckets; ++q) {
        const int b = kNumBuckets - kBigBuckets + q;
        bl.list[q] = lists[b];
        bl.count[q] = counts[b];
        bl.off[q] = listed;
        listed += counts[b];
        items += counts[b] * kBigParts[q];
    }
    if (!listed) return;
    if (!a.counters_zeroed) fill_u32(a.n_movers, 0u, kBigBuckets, s);
    hipLaunchKernelGGL(k_big_movers, dim3((listed + kBlock - 1) / kBlock), blk, 0, s, bl, a.comm,
                       (const uint32_t *)a.target, a.movers, a.n_movers);
    const dim3 grid(std::min<uint32_t>(items, kCommitBlocks));
    if (ident) {
        if (wtd) hipLaunchKernelGGL((k_commit_big<true, kCommitIdent>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
        else hipLaunchKernelGGL((k_commit_big<false, kCommitIdent>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
    } else {
        if (bit) {
            if (wtd) hipLaunchKernelGGL((k_commit_big<true, kCommitPackedBit>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
            else hipLaunchKernelGGL((k_commit_big<false, kCommitPackedBit>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
        } else {
            if (wtd) hipLaunchKernelGGL((k_commit_big<true, kCommitPacked>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
            else hipLaunchKernelGGL((k_commit_big<false, kCommitPacked>), grid, blk, 0, s, c, bl, a.movers, a.n_movers);
        }
    }
}

void finalize_pass(const SweepArgs &a, uint32_t *const *next_lists, uint32_t *next_counts,
                   hipStream_t s) {
    ListPtrs lp;
    for (int b = 0; b < kNumBuckets; ++b) lp.p[b] = next_lists[b];
    if (!a.counters_zeroed) fill_u32(next_counts, 0u, kNumBuckets, s);
    const uint32_t n = a.g.n;
    if (!n) return;
    const uint32_t per = (uint32_t)(kBlock * kFinTiles);
    hipLaunchKernelGGL(k_finalize, dim3((n + per - 1) / per), dim3(kBlock), 0, s,
                       const_cast<uint32_t *>(a.comm), (const uint32_t *)a.target,
                       (const uint32_t *)a.tot, a.mark, a.bucket, n, lp, next_counts, a.stats,
                       a.bt.gid, a.bt.sumsq, marks_in_ct(a) ? a.ct : nullptr);
}

void renumber_flags(const uint32_t *tot, uint32_t n, uint32_t *flags, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_renumber_flags, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, tot, n,
                       flags);
}

void update_best(uint32_t *best, uint32_t n0, const uint32_t *comm, const uint32_t *map,
                 hipStream_t s) {
    if (!n0) return;
    hipLaunchKernelGGL(k_update_best, dim3(grid_for(n0, kBlock * 4)), dim3(kBlock), 0, s, best, n0,
                       comm, map);
}

void new_degrees(const uint32_t *tot, const uint32_t *map, uint32_t n, uint32_t *k_new,
                 hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_new_degrees, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, tot, map,
                       n, k_new);
}

}  // namespace lv
