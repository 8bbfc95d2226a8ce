// lv_aggregate.hip -- contraction of a level into the graph of its communities
// (reference: AggregationPhase::run, phaseAggregation.cuh:51-170: relabel every
// entry to (new(comm[src]), new(comm[dst])), merge parallel entries by summing
// their weights, rebuild the CSR, degrees and offsets).
//
// Pipeline (all on the device, integer weights, exact):
//   1. newc[v] = new id of v's community; vertices sorted by newc (radix sort)
//      give the member list of every community and P, the prefix of member
//      degrees, so community c owns the "virtual entries" [crow[c], crow[c+1]).
//   2. Communities are bucketed by crow[c+1]-crow[c] like vertices in the
//      best-move pass.  A group of G lanes (up to a workgroup) stages member
//      offsets in LDS, accumulates the weight towards every destination
//      community in an LDS table, then writes the distinct (dst, w) pairs to a
//      staging area at crow[c].
//   3. Giant communities (above 4096 virtual entries) gather their entries as
//      keys (community << bits | dst), sort them with the LSD radix sort and
//      reduce runs of equal keys.
//   4. Row sizes are scanned into the new offsets and every row is copied to
//      place.  Rows of LDS-table communities stay in table order: nothing
//      downstream depends on the order of a row; exported graphs are sorted
//      by destination on the way out (the reference's sort_by_key order).
#include "lv_device.h"

namespace lv {

namespace {

constexpr int kAggIlp = 4;   // entries per lane with loads in flight together

struct AggCtx {
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *wt;        // nullptr: unit weights
    const uint32_t *newc;      // [n]
    const uint32_t *members;   // [n] vertices sorted by new community
    const uint32_t *mstart;    // [n_new + 1]
    const uint64_t *P;         // [n + 1] prefix of member degrees
    const uint64_t *crow;      // [n_new + 1]
    uint32_t *tcol;            // [E] staging: distinct destinations
    uint32_t *tw;              // [E] staging: their weights
    uint32_t *rcount;          // [n_new] distinct destinations per community
};

// Barrier over the G lanes that share one community: a wavefront barrier with
// LDS ordering for groups inside one wave, the workgroup barrier otherwise.
template <int G>
__device__ inline void group_barrier() {
    if constexpr (G == kBlock) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Largest i in [0, n) with p[i] <= k, for a staged LDS prefix (p[0] <= k).
__device__ inline uint32_t lds_member(const uint64_t *p, uint32_t n, uint64_t k) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p[mid] <= k) lo = mid; else hi = mid;
    }
    return lo;
}

// Largest i in [lo, hi) with P[i] <= k (P[lo] <= k < P[hi] or hi is the end).
__device__ inline uint32_t member_of(const uint64_t *P, uint32_t lo, uint32_t hi, uint64_t k) {
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (P[mid] <= k) lo = mid; else hi = mid;
    }
    return lo;
}

template <int G, int T, bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_agg_table(AggCtx a, const uint32_t *__restrict__ list,
                                                      uint32_t count) {
    constexpr int GPB = kBlock / G;
    __shared__ uint32_t s_key[GPB * T];
    __shared__ uint32_t s_val[GPB * T];
    __shared__ uint32_t s_cnt[GPB];
    __shared__ uint32_t s_pos[GPB];
    const int lane = threadIdx.x & (G - 1);
    const int gi = threadIdx.x / G;
    const uint32_t idx = blockIdx.x * GPB + gi;
    const bool valid = idx < count;
    const uint32_t c = valid ? list[idx] : 0u;
    uint32_t *K = s_key + gi * T;
    uint32_t *V = s_val + gi * T;
    uint64_t k0 = 0;
    uint32_t lg = 1;
    if (valid) {
        k0 = a.crow[c];
        lg = table_log2(a.crow[c + 1] - k0);   // at most 2x the virtual entries, <= T
    }
    const uint32_t tv = 1u << lg, mask = tv - 1u;
    for (uint32_t s = lane; s < tv; s += G) { K[s] = LV_EMPTY; V[s] = 0u; }
    if (lane == 0) { s_cnt[gi] = 0u; s_pos[gi] = 0u; }
    __syncthreads();
    // members staged G at a time: row start and first virtual entry in LDS
    __shared__ uint64_t s_mrow[GPB][G];
    __shared__ uint64_t s_mp[GPB][G + 1];
    const uint32_t m0 = valid ? a.mstart[c] : 0u;
    const uint32_t m1 = valid ? a.mstart[c + 1] : 0u;
    for (uint32_t mb = m0; mb < m1; mb += G) {
        const uint32_t nm = m1 - mb < (uint32_t)G ? m1 - mb : (uint32_t)G;
        if ((uint32_t)lane < nm) {
            const uint32_t v = a.members[mb + lane];
            s_mrow[gi][lane] = a.row[v];
            s_mp[gi][lane] = a.P[mb + lane];
        }
        if (lane == 0) s_mp[gi][nm] = a.P[mb + nm];
        group_barrier<G>();
        const uint64_t e0 = s_mp[gi][0], e1 = s_mp[gi][nm];
        for (uint64_t kb = e0 + lane; kb < e1; kb += (uint64_t)G * kAggIlp) {
            uint32_t xs[kAggIlp], ws[kAggIlp], ds[kAggIlp];
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                const uint64_t k = kb + (uint64_t)u * G;
                xs[u] = LV_EMPTY;
                ws[u] = 1u;
                if (k < e1) {
                    const uint32_t i = lds_member(s_mp[gi], nm, k);
                    const uint64_t j = s_mrow[gi][i] + (k - s_mp[gi][i]);
                    xs[u] = stream_load(&a.col[j]);
                    if (WEIGHTED) ws[u] = stream_load(&a.wt[j]);
                }
            }
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) ds[u] = xs[u] != LV_EMPTY ? a.newc[xs[u]] : LV_EMPTY;
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                const uint32_t d = ds[u];
                if (d == LV_EMPTY) continue;
                uint32_t h = slot_hash_dyn(d, lg);
                for (;;) {
                    const uint32_t prev = atomicCAS(&K[h], LV_EMPTY, d);
                    if (prev == LV_EMPTY) { atomicAdd(&s_cnt[gi], 1u); atomicAdd(&V[h], ws[u]); break; }
                    if (prev == d) { atomicAdd(&V[h], ws[u]); break; }
                    h = (h + 1) & mask;
                }
            }
        }
        group_barrier<G>();
    }
    __syncthreads();
    // distinct (destination, weight) pairs in table order; the row order is not
    // part of the result (every consumer is order independent; exported graphs
    // are put in destination order on the way out)
    if (valid) {
        if (lane == 0) a.rcount[c] = s_cnt[gi];
        for (uint32_t s = lane; s < tv; s += G) {
            const uint32_t d = K[s];
            if (d == LV_EMPTY) continue;
            const uint32_t pos = atomicAdd(&s_pos[gi], 1u);
            a.tcol[k0 + pos] = d;
            a.tw[k0 + pos] = V[s];
        }
    }
}

struct AggHub {
    AggCtx a;
    const uint64_t *hoff;      // [hub communities + 1] offsets of their entries
    uint32_t parts;
    int bits;                  // destination bits
    uint64_t *hkey;            // (hub index << bits) | destination
    uint32_t *hval;            // weight
};

// Gather the virtual entries of the giant communities as sort keys.
template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_agg_hub_gather(AggHub h, const uint32_t *__restrict__ list,
                                                           uint32_t count) {
    const uint32_t ci = blockIdx.x / h.parts;
    const uint32_t part = blockIdx.x % h.parts;
    if (ci >= count) return;
    const AggCtx &a = h.a;
    const uint32_t c = list[ci];
    const uint64_t k0 = a.crow[c], k1 = a.crow[c + 1];
    const uint64_t span = (k1 - k0 + h.parts - 1) / h.parts;
    const uint64_t s0 = k0 + span * part;
    const uint64_t s1 = s0 + span < k1 ? s0 + span : k1;
    const uint32_t m0 = a.mstart[c], m1 = a.mstart[c + 1];
    const uint64_t out = h.hoff[ci] - k0;
    const uint64_t hi_bits = (uint64_t)ci << h.bits;
    __shared__ uint64_t s_mrow[kBlock];
    __shared__ uint64_t s_mp[kBlock + 1];
    __shared__ uint32_t s_first;
    if (threadIdx.x == 0) s_first = s0 < s1 ? member_of(a.P, m0, m1, s0) : m1;
    __syncthreads();
    for (uint32_t mb = s_first; mb < m1; mb += kBlock) {
        const uint32_t nm = m1 - mb < (uint32_t)kBlock ? m1 - mb : (uint32_t)kBlock;
        if (threadIdx.x < nm) {
            const uint32_t v = a.members[mb + threadIdx.x];
            s_mrow[threadIdx.x] = a.row[v];
            s_mp[threadIdx.x] = a.P[mb + threadIdx.x];
        }
        if (threadIdx.x == 0) s_mp[nm] = a.P[mb + nm];
        __syncthreads();
        const uint64_t e0 = s_mp[0] > s0 ? s_mp[0] : s0;
        const uint64_t e1 = s_mp[nm] < s1 ? s_mp[nm] : s1;
        for (uint64_t k = e0 + threadIdx.x; k < e1; k += kBlock) {
            const uint32_t i = lds_member(s_mp, nm, k);
            const uint64_t j = s_mrow[i] + (k - s_mp[i]);
            h.hkey[out + k] = hi_bits | a.newc[a.col[j]];
            h.hval[out + k] = WEIGHTED ? a.wt[j] : 1u;
        }
        const bool done = s_mp[nm] >= s1;
        __syncthreads();
        if (done) break;
    }
}

// Runs of equal keys in the sorted hub entries: head flags.
__global__ void k_heads(const uint64_t *keys, uint64_t n, uint32_t *heads) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x)
        heads[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}

// With idx = inclusive scan of heads and S = inclusive scan of weights:
// unique key, and weight of each run as S[last] - S[first] + w[first].
__global__ void k_runs(const uint64_t *keys, const uint32_t *vals, const uint32_t *idx,
                       const uint64_t *S, uint64_t n, uint64_t *ukey, uint64_t *ws, uint64_t *we) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = idx[j] - 1u;
        if (j == 0 || keys[j] != keys[j - 1]) {
            ukey[u] = keys[j];
            ws[u] = S[j] - vals[j];
        }
        if (j + 1 == n || keys[j + 1] != keys[j]) we[u] = S[j];
    }
}

// Start of each hub community's row among the unique keys.
__global__ void k_hub_starts(const uint64_t *ukey, uint64_t nu, int bits, uint32_t *ustart) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < nu;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t hc = (uint32_t)(ukey[u] >> bits);
        if (u == 0 || (uint32_t)(ukey[u - 1] >> bits) != hc) ustart[hc] = (uint32_t)u;
    }
}

__global__ void k_hub_counts(const uint32_t *ustart, const uint32_t *list, uint32_t count,
                             uint32_t *rcount) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        rcount[list[i]] = ustart[i + 1] - ustart[i];
}

// Copy every community's sorted row to its final offset.  Table-bucket rows
// come from the staging area; giant communities from the unique hub keys.
template <int G, bool HUB>
__global__ __launch_bounds__(kBlock) void k_agg_copy(AggCtx a, const uint32_t *__restrict__ list,
                                                     uint32_t count, uint32_t parts,
                                                     const uint64_t *out_row, uint32_t *col,
                                                     uint32_t *w, const uint32_t *ustart,
                                                     const uint64_t *ukey, const uint64_t *ws,
                                                     const uint64_t *we, uint32_t mask) {
    constexpr int GPB = kBlock / G;
    const int lane = threadIdx.x & (G - 1);
    uint32_t idx, part;
    if constexpr (G == kBlock) {
        idx = blockIdx.x / parts;
        part = blockIdx.x % parts;
    } else {
        idx = blockIdx.x * GPB + threadIdx.x / G;
        part = 0;
    }
    if (idx >= count) return;
    const uint32_t c = list[idx];
    const uint32_t r = a.rcount[c];
    const uint64_t dst = out_row[c];
    const uint32_t stride = G * (G == kBlock ? parts : 1u);
    if constexpr (HUB) {
        const uint64_t src = ustart[idx];
        for (uint32_t t = part * G + lane; t < r; t += stride) {
            col[dst + t] = (uint32_t)ukey[src + t] & mask;
            w[dst + t] = (uint32_t)(we[src + t] - ws[src + t]);
        }
    } else {
        const uint64_t src = a.crow[c];
        for (uint32_t t = part * G + lane; t < r; t += stride) {
            col[dst + t] = a.tcol[src + t];
            w[dst + t] = a.tw[src + t];
        }
    }
}

__global__ void k_newc(const uint32_t *comm, const uint32_t *map, uint32_t n, uint32_t *newc,
                       uint64_t *key, uint32_t *val) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = renumbered(map, comm[v]);
        newc[v] = c;
        key[v] = c;
        val[v] = (uint32_t)v;
    }
}

// After the vertex sort: member starts, member degrees.
__global__ void k_members(const uint64_t *skey, const uint32_t *members, const uint64_t *row,
                          uint32_t n, uint32_t *mstart, uint64_t *mdeg) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        if (i == 0 || skey[i] != skey[i - 1]) mstart[skey[i]] = (uint32_t)i;
        const uint32_t v = members[i];
        mdeg[i] = row[v + 1] - row[v];
    }
}

__global__ void k_comm_rows(const uint64_t *P, const uint32_t *mstart, uint32_t n_new,
                            uint64_t *crow) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c <= n_new;
         c += (size_t)gridDim.x * blockDim.x)
        crow[c] = P[mstart[c]];
}

__global__ void k_widen(const uint32_t *in, uint32_t n, uint64_t *out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }
__global__ void k_set_u64(uint64_t *p, uint64_t v) { *p = v; }

}  // namespace

void agg_newc(const uint32_t *comm, const uint32_t *map, uint32_t n, uint32_t *newc, uint64_t *key,
              uint32_t *val, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_newc, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, comm, map, n,
                       newc, key, val);
}

void agg_members(const uint64_t *skey, const uint32_t *members, const uint64_t *row, uint32_t n,
                 uint32_t n_new, uint32_t *mstart, uint64_t *mdeg, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_members, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, skey, members,
                       row, n, mstart, mdeg);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, mstart + n_new, n);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, mdeg + n, (uint64_t)0);
}

void agg_comm_rows(const uint64_t *P, const uint32_t *mstart, uint32_t n_new, uint64_t *crow,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_comm_rows, dim3(grid_for((size_t)n_new + 1, kBlock * 4)), dim3(kBlock), 0,
                       s, P, mstart, n_new, crow);
}

void set_u64(uint64_t *p, uint64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, p, v);
}

void widen_u32(const uint32_t *in, uint32_t n, uint64_t *out, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_widen, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, in, n, out);
}

static AggCtx ctx_of(const AggArgs &x) {
    AggCtx a;
    a.row = x.g.row; a.col = x.g.col; a.wt = x.g.wt;
    a.newc = x.newc; a.members = x.members; a.mstart = x.mstart; a.P = x.P; a.crow = x.crow;
    a.tcol = x.tcol; a.tw = x.tw; a.rcount = x.rcount;
    return a;
}

void agg_tables(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts, hipStream_t s) {
    const AggCtx a = ctx_of(x);
    const bool wtd = x.g.wt != nullptr;
    for (int b = 0; b < kHubBucket; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        const uint32_t *L = lists[b];
#define LV_AGG(GG, TT)                                                                          \
    do {                                                                                        \
        const dim3 grid((cnt + kBlock / (GG) - 1) / (kBlock / (GG)));                           \
        if (wtd) hipLaunchKernelGGL((k_agg_table<GG, TT, true>), grid, dim3(kBlock), 0, s, a, L, cnt); \
        else hipLaunchKernelGGL((k_agg_table<GG, TT, false>), grid, dim3(kBlock), 0, s, a, L, cnt); \
    } while (0)
        switch (b) {
            case 0: LV_AGG(4, 8); break;
            case 1: LV_AGG(8, 16); break;
            case 2: LV_AGG(16, 32); break;
            case 3: LV_AGG(32, 64); break;
            case 4: LV_AGG(64, 128); break;
            case 5: LV_AGG(64, 512); break;
            case 6: LV_AGG(256, 2048); break;
            default: LV_AGG(256, 8192); break;
        }
#undef LV_AGG
    }
}

void agg_hub_gather(const AggArgs &x, const uint32_t *hub_list, uint32_t nh, const uint64_t *hoff,
                    uint32_t parts, int bits, uint64_t *hkey, uint32_t *hval, hipStream_t s) {
    if (!nh) return;
    AggHub h;
    h.a = ctx_of(x);
    h.hoff = hoff;
    h.parts = parts;
    h.bits = bits;
    h.hkey = hkey;
    h.hval = hval;
    const dim3 grid(nh * parts);
    if (x.g.wt) hipLaunchKernelGGL(k_agg_hub_gather<true>, grid, dim3(kBlock), 0, s, h, hub_list, nh);
    else hipLaunchKernelGGL(k_agg_hub_gather<false>, grid, dim3(kBlock), 0, s, h, hub_list, nh);
}

void heads_u64(const uint64_t *keys, uint64_t n, uint32_t *heads, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_heads, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, keys, n, heads);
}

void agg_runs(const uint64_t *keys, const uint32_t *vals, const uint32_t *idx, const uint64_t *S,
              uint64_t n, uint64_t *ukey, uint64_t *ws, uint64_t *we, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_runs, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, keys, vals, idx, S,
                       n, ukey, ws, we);
}

void agg_hub_rows(const uint64_t *ukey, uint64_t nu, int bits, const uint32_t *hub_list,
                  uint32_t nh, uint32_t *ustart, uint32_t *rcount, hipStream_t s) {
    if (!nh) return;
    hipLaunchKernelGGL(k_hub_starts, dim3(grid_for(nu, kBlock * 4)), dim3(kBlock), 0, s, ukey, nu,
                       bits, ustart);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, ustart + nh, (uint32_t)nu);
    hipLaunchKernelGGL(k_hub_counts, dim3(grid_for(nh, kBlock)), dim3(kBlock), 0, s,
                       (const uint32_t *)ustart, hub_list, nh, rcount);
}

void agg_copy(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts, uint32_t hub_parts,
              const uint64_t *out_row, uint32_t *col, uint32_t *w, const uint32_t *ustart,
              const uint64_t *ukey, const uint64_t *ws, const uint64_t *we, uint32_t mask,
              hipStream_t s) {
    const AggCtx a = ctx_of(x);
    for (int b = 0; b < kNumBuckets; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        const uint32_t *L = lists[b];
#define LV_CP(GG, PARTS, HUB)                                                                   \
    do {                                                                                        \
        const uint32_t gpb = kBlock / (GG);                                                     \
        const dim3 grid((GG) == kBlock ? cnt * (PARTS) : (cnt + gpb - 1) / gpb);                \
        hipLaunchKernelGGL((k_agg_copy<GG, HUB>), grid, dim3(kBlock), 0, s, a, L, cnt,          \
                           (uint32_t)(PARTS), out_row, col, w, ustart, ukey, ws, we, mask);     \
    } while (0)
        switch (b) {
            case 0: LV_CP(4, 1, false); break;
            case 1: LV_CP(8, 1, false); break;
            case 2: LV_CP(16, 1, false); break;
            case 3: LV_CP(32, 1, false); break;
            case 4: LV_CP(64, 1, false); break;
            case 5: LV_CP(64, 1, false); break;
            case 6: LV_CP(256, 1, false); break;
            case 7: LV_CP(256, 4, false); break;
            default: LV_CP(256, hub_parts, true); break;
        }
#undef LV_CP
    }
}

}  // namespace lv
