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
//      (dst, w) pairs, move them into destination hash partitions and sum
//      each partition per destination in an LDS table.
//   4. Row sizes are scanned into the new offsets and every row is copied to
//      place.  Rows of LDS-table communities stay in table order: nothing
//      downstream depends on the order of a row; exported graphs are sorted
//      by destination on the way out (the reference's sort_by_key order).
#include "lv_device.h"

namespace lv {

namespace {

constexpr int kAggIlp = 4;   // entries per lane with loads in flight together
constexpr int kCopyIlp = 8;  // row copies: loads per lane issued before their stores

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
    // the next level's first sweep decided from the rows (DEC kernels)
    int64_t m2;
    double tw_total;
    const unsigned long long *newck;   // [n] newc | new degree << 32
    const uint32_t *knew;              // [n_new] new degrees
    uint32_t *dtarget;                 // [n_new]
    unsigned long long *dstats;        // one statistics section
    BatchView dbt;                     // batches: graph, 2m, m and live flag per new vertex
};

// The next level's first sweep, decided from a contracted row: that sweep is
// a fast sweep at the identity partition (phaseOptimizationSort.cu:28-158 with
// the v < c guard of phaseOptimizationUpdate.cu:99), so new vertex c's
// candidates are the distinct destinations d != c of its row, each with the
// row's summed weight w_cd and the total of d's singleton community, d's new
// degree k_d; c's own total is k_c.  DEC = 1 + the gain mode.  The same
// scores, order and move rule as the fast-sweep kernels (k_move_fast); in a
// batch each vertex takes its graph's 2m and m, and a vertex of a graph that
// does not sweep the next level is not examined (vertex_graph).
template <int DEC>
struct FirstSweep {
    using SC = Score<DEC - 1>;
    using S = typename SC::T;
    __device__ static S score(uint32_t w, uint32_t kc, uint32_t kd, int64_t m2, double tw) {
        return SC::of(w, 0u, kc, kc, kd, m2, tw);
    }
    // the decision for c from its best candidate
    __device__ static uint32_t decide(uint32_t c, uint32_t kc, S best, uint32_t best_c,
                                      int64_t m2) {
        const bool mv = best != SC::none() && SC::positive(best, 0u, kc, kc, m2) && c < best_c;
        return mv ? best_c : c;
    }
};

// Largest i in [i0, n) with p[i] <= k, for a staged LDS prefix with
// p[i0] <= k: a gallop forward from i0, then a binary search in the bracket.
// A lane walks its entries in increasing order, so i0 (the member of its
// previous entry) is at most a few members behind: one to three LDS reads
// instead of log2(n).
__device__ inline uint32_t lds_member_from(const uint64_t *p, uint32_t n, uint64_t k, uint32_t i0) {
    uint32_t lo = i0, hi = n, step = 1;
    while (lo + step < n) {
        if (p[lo + step] <= k) {
            lo += step;
            step <<= 1;
        } else {
            hi = lo + step;
            break;
        }
    }
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

// DEC = 0: the rows only.  DEC = 1 + gain: the rows and the next level's
// first sweep (FirstSweep); tables of at most kPackedMaxT slots then gather
// newck[x] (destination and its new degree in one word) and keep each
// destination's degree beside it, larger ones gather the degree per
// destination when they score.
template <int G, int T, bool WEIGHTED, int DEC>
__global__ __launch_bounds__(kBlock) void k_agg_table(AggCtx a, const uint32_t *__restrict__ list,
                                                      uint32_t count) {
    constexpr int GPB = kBlock / G;
    constexpr bool PK = DEC != 0 && T <= kPackedMaxT;
    constexpr int TK = PK ? T : 1;
    // members staged at a time: G, or fewer where the staging would
    // otherwise cost a workgroup per CU (bucket 5 and 6 tables at 8 per CU
    // with their degree arrays, bucket 7's at 4)
    constexpr int SM = T == kTable7 ? 64 : !PK ? G : (G >= 256 ? 64 : (G >= 64 ? 16 : G));
    __shared__ uint32_t s_key[GPB * T];
    __shared__ uint32_t s_val[GPB * T];
    __shared__ uint32_t s_kd[GPB * TK];   // PK: new degree of each destination
    __shared__ uint32_t s_cnt[GPB];
    __shared__ uint32_t s_pos[GPB];
    const int lane = threadIdx.x & (G - 1);
    const int gi = (int)group_index<G>();
    const uint32_t idx = blockIdx.x * GPB + gi;
    const bool valid = idx < count;
    const uint32_t c = valid ? list[idx] : 0u;
    uint32_t *K = s_key + gi * T;
    uint32_t *V = s_val + gi * T;
    uint32_t *KD = s_kd + gi * TK;
    uint64_t k0 = 0;
    uint32_t tv = 4u;
    // the community's words together: row bounds and member range
    const uint32_t m0 = valid ? a.mstart[c] : 0u;
    const uint32_t m1 = valid ? a.mstart[c + 1] : 0u;
    if (valid) {
        k0 = a.crow[c];
        // at most two thirds full with an empty slot (the virtual entries
        // bound the distinct destinations), or T (bucket 7: above every
        // entry count of the bucket)
        tv = min(table_slots(a.crow[c + 1] - k0), (uint32_t)T);
    }
    for (uint32_t s = lane; s < tv; s += G) { K[s] = LV_EMPTY; V[s] = 0u; }
    if (lane == 0) { s_cnt[gi] = 0u; s_pos[gi] = 0u; }
    // a community's table is its group's alone
    group_barrier<G>();
    // members staged SM at a time: row start and first virtual entry in LDS
    __shared__ uint64_t s_mrow[GPB][SM];
    __shared__ uint64_t s_mp[GPB][SM + 1];
    for (uint32_t mb = m0; mb < m1; mb += SM) {
        const uint32_t nm = m1 - mb < (uint32_t)SM ? m1 - mb : (uint32_t)SM;
        if ((uint32_t)lane < nm) {
            const uint32_t v = a.members[mb + lane];
            s_mrow[gi][lane] = a.row[v];
            s_mp[gi][lane] = a.P[mb + lane];
        }
        if (lane == 0) s_mp[gi][nm] = a.P[mb + nm];
        group_barrier<G>();
        const uint64_t e0 = s_mp[gi][0], e1 = s_mp[gi][nm];
        uint32_t mi = 0;   // member of the lane's previous entry
        for (uint64_t kb = e0 + lane; kb < e1; kb += (uint64_t)G * kAggIlp) {
            uint32_t xs[kAggIlp], ws[kAggIlp], ds[kAggIlp], kds[kAggIlp];
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                const uint64_t k = kb + (uint64_t)u * G;
                xs[u] = LV_EMPTY;
                ws[u] = 1u;
                if (k < e1) {
                    const uint32_t i = lds_member_from(s_mp[gi], nm, k, mi);
                    mi = i;
                    const uint64_t j = s_mrow[gi][i] + (k - s_mp[gi][i]);
                    xs[u] = stream_load(&a.col[j]);
                    if (WEIGHTED) ws[u] = stream_load(&a.wt[j]);
                }
            }
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                if constexpr (PK) {
                    const unsigned long long p = xs[u] != LV_EMPTY ? a.newck[xs[u]] : ~0ull;
                    ds[u] = (uint32_t)p;
                    kds[u] = (uint32_t)(p >> 32);
                } else {
                    ds[u] = xs[u] != LV_EMPTY ? a.newc[xs[u]] : LV_EMPTY;
                    kds[u] = 0u;
                }
            }
            // first probes issued back to back (one LDS round trip), then the
            // rare occupied first slots probe on; a lane's entries stay in order
            uint32_t hs[kAggIlp], pv[kAggIlp];
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                hs[u] = slot_hash_mod(ds[u], tv);
                pv[u] = ds[u] != LV_EMPTY ? atomicCAS(&K[hs[u]], LV_EMPTY, ds[u]) : LV_EMPTY;
            }
#pragma unroll
            for (int u = 0; u < kAggIlp; ++u) {
                const uint32_t d = ds[u];
                if (d == LV_EMPTY) continue;
                uint32_t h = hs[u], prev = pv[u];
                while (prev != LV_EMPTY && prev != d) {
                    h = h + 1 == tv ? 0u : h + 1;
                    prev = atomicCAS(&K[h], LV_EMPTY, d);
                }
                if (prev == LV_EMPTY) {
                    atomicAdd(&s_cnt[gi], 1u);
                    if constexpr (PK) KD[h] = kds[u];
                }
                atomicAdd(&V[h], ws[u]);
            }
        }
        group_barrier<G>();
    }
    group_barrier<G>();
    // distinct (destination, weight) pairs in table order; the row order is not
    // part of the result (every consumer is order independent; exported graphs
    // are put in destination order on the way out)
    if constexpr (DEC == 0) {
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
    } else {
        using FS = FirstSweep<DEC>;
        using S = typename FS::S;
        S best = FS::SC::none();
        uint32_t best_c = LV_EMPTY;
        uint32_t nself = 0;
        const uint32_t kc = valid ? a.knew[c] : 0u;
        int64_t m2 = a.m2;
        double tw = a.tw_total;
        const bool live = valid && vertex_graph(a.dbt, c, a.m2, a.tw_total, m2, tw);
        if (valid) {
            if (lane == 0) a.rcount[c] = s_cnt[gi];
            // kAggIlp slots per lane at a time: without the packed degrees
            // their destinations' degrees are gathered together
            for (uint32_t sb = lane; sb < tv; sb += (uint32_t)G * kAggIlp) {
                uint32_t ds[kAggIlp], ws[kAggIlp], kds[kAggIlp];
#pragma unroll
                for (int u = 0; u < kAggIlp; ++u) {
                    const uint32_t sl = sb + (uint32_t)u * G;
                    ds[u] = sl < tv ? K[sl] : LV_EMPTY;
                    ws[u] = sl < tv ? V[sl] : 0u;
                    if constexpr (PK) kds[u] = sl < tv ? KD[sl] : 0u;
                }
                if constexpr (!PK) {
#pragma unroll
                    for (int u = 0; u < kAggIlp; ++u)
                        kds[u] = live && ds[u] != LV_EMPTY && ds[u] != c ? a.knew[ds[u]] : 0u;
                }
#pragma unroll
                for (int u = 0; u < kAggIlp; ++u) {
                    const uint32_t d = ds[u];
                    if (d == LV_EMPTY) continue;
                    const uint32_t pos = atomicAdd(&s_pos[gi], 1u);
                    a.tcol[k0 + pos] = d;
                    a.tw[k0 + pos] = ws[u];
                    if (d == c || !live) continue;   // the self-loop is no candidate
                    ++nself;
                    const S sc = FS::score(ws[u], kc, kds[u], m2, tw);
                    if (better_candidate(sc, d, best, best_c, false)) { best = sc; best_c = d; }
                }
            }
        }
        uint32_t considered = 0, twrite = 0;
        if constexpr (G <= 64) {
            group_best<G, S>(best, best_c, false);
            const bool any = group_any<G>(nself > 0);
            if (valid && lane == 0) {
                considered = any;
                const uint32_t t = live ? FS::decide(c, kc, best, best_c, m2) : c;
                a.dtarget[c] = t;
                twrite = t != c;
            }
        } else {
            __shared__ S r_s[kBlock / 64];
            __shared__ uint32_t r_c[kBlock / 64];
            __shared__ uint32_t r_n[kBlock / 64];
            group_best<64, S>(best, best_c, false);
            const uint32_t wn = group_any<64>(nself > 0);
            const int w = threadIdx.x >> 6;
            if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; r_n[w] = wn; }
            __syncthreads();
            if (threadIdx.x == 0 && valid) {
                uint32_t tn = 0;
                for (int i = 0; i < kBlock / 64; ++i) {
                    tn += r_n[i];
                    if (better_candidate(r_s[i], r_c[i], best, best_c, false)) {
                        best = r_s[i];
                        best_c = r_c[i];
                    }
                }
                considered = tn > 0;
                const uint32_t t = live ? FS::decide(c, kc, best, best_c, m2) : c;
                a.dtarget[c] = t;
                twrite = t != c;
            }
        }
        block_add_stats(a.dstats, ST_E_ACTIVE, nself, ST_CANDIDATES, nself, ST_CONSIDERED,
                        considered, ST_TWRITES, twrite);
    }
}

// ---- giant communities (above 4096 virtual entries) ----
//
// A giant community's destinations are cut into 2^lgp partitions by the top
// bits of hub_mix (one per kGiantPartEntries entries, at most 4096) and its
// virtual entries into tiles of 2^lgt (at least 16 entries per partition, so
// a tile reserves at most one run per 16 entries).  k_giant_gather writes the
// (destination, weight) pairs of a tile, packed at the start of the tile's
// range, folding repeats of the destinations it caches (a community's own id
// and the giants it is bound to repeat hundreds of thousands of times) into
// one pair per tile, and reserves each (tile, partition) run; with poff =
// exclusive scan of the partition counts (every pair is in one partition, so
// poff is the pair buffers' position of a partition), k_giant_scatter moves
// a tile's pairs to partition order;
// k_giant_reduce sums one partition per workgroup in the LDS table
// (lv_device.h) and writes its distinct pairs back over the tile-order buffer;
// k_giant_copy puts them at the community's final row.

constexpr int kGiantIlp = 8;
constexpr uint32_t kGiantCache = 1024;                 // destinations cached per tile

struct GiantCtx : GiantArgs {
    AggCtx a;
};

__device__ inline uint32_t giant_part(uint32_t d, uint32_t lgp) {
    return lgp ? hub_mix(d) >> (32 - lgp) : 0u;
}

template <bool WEIGHTED>
__global__ __launch_bounds__(kBlock) void k_giant_gather(GiantCtx h) {
    extern __shared__ uint32_t hist[];   // [h.max_parts]
    __shared__ uint64_t s_mrow[kBlock];
    __shared__ uint64_t s_mp[kBlock + 1];
    __shared__ uint32_t s_first;
    __shared__ uint32_t s_n;             // pairs written so far
    // destinations cached for the tile: a repeat adds its weight; each cached
    // destination is written once, after the tile's other pairs
    __shared__ uint32_t c_key[kGiantCache];
    __shared__ uint32_t c_val[kGiantCache];
    const AggCtx &a = h.a;
    const uint64_t tile = h.tiles[blockIdx.x];
    const uint32_t lt = (uint32_t)tile;
    const GiantInfo I = h.info[tile >> 32];
    const uint32_t np = 1u << I.lgp;
    for (uint32_t i = threadIdx.x; i < np; i += kBlock) hist[i] = 0u;
    for (uint32_t i = threadIdx.x; i < kGiantCache; i += kBlock) { c_key[i] = LV_EMPTY; c_val[i] = 0u; }
    const uint64_t k0 = a.crow[I.c], k1 = a.crow[I.c + 1];
    const uint64_t s0 = k0 + ((uint64_t)lt << I.lgt);
    const uint64_t s1 = s0 + (1ull << I.lgt) < k1 ? s0 + (1ull << I.lgt) : k1;
    const uint32_t m0 = a.mstart[I.c], m1 = a.mstart[I.c + 1];
    // the tile's range of the pair buffer: at most one pair per entry
    unsigned long long *out = h.p0 + I.off + (s0 - k0);
    if (threadIdx.x == 0) {
        s_first = member_of(a.P, m0, m1, s0);
        s_n = 0u;
        if (lt == 0) a.rcount[I.c] = 0u;   // the reduce adds the distinct counts
    }
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
        uint32_t mi = 0;   // member of the lane's previous entry
        for (uint64_t kb = e0 + threadIdx.x; kb < e1; kb += (uint64_t)kBlock * kGiantIlp) {
            uint32_t xs[kGiantIlp], ws[kGiantIlp], ds[kGiantIlp];
#pragma unroll
            for (int u = 0; u < kGiantIlp; ++u) {
                const uint64_t k = kb + (uint64_t)u * kBlock;
                xs[u] = LV_EMPTY;
                ws[u] = 1u;
                if (k < e1) {
                    const uint32_t i = lds_member_from(s_mp, nm, k, mi);
                    mi = i;
                    const uint64_t j = s_mrow[i] + (k - s_mp[i]);
                    xs[u] = stream_load(&a.col[j]);
                    if (WEIGHTED) ws[u] = stream_load(&a.wt[j]);
                }
            }
#pragma unroll
            for (int u = 0; u < kGiantIlp; ++u) ds[u] = xs[u] != LV_EMPTY ? a.newc[xs[u]] : LV_EMPTY;
#pragma unroll
            for (int u = 0; u < kGiantIlp; ++u) {
                const uint32_t d = ds[u];
                bool emit = false;
                if (d != LV_EMPTY) {
                    const uint32_t cs = hub_mix(d) & (kGiantCache - 1);
                    uint32_t cur = c_key[cs];
                    if (cur == LV_EMPTY) cur = atomicCAS(&c_key[cs], LV_EMPTY, d);
                    if (cur == LV_EMPTY || cur == d) {   // cached here, or a repeat
                        atomicAdd(&c_val[cs], ws[u]);
                    } else {
                        emit = true;
                        atomicAdd(&hist[giant_part(d, I.lgp)], 1u);
                    }
                }
                wave_append(emit, ((unsigned long long)d << 32) | ws[u], &s_n, out);
            }
        }
        const bool done = s_mp[nm] >= s1;
        __syncthreads();
        if (done) break;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kGiantCache; i += kBlock) {   // uniform trip count
        const uint32_t d = c_key[i];
        const bool emit = d != LV_EMPTY;
        if (emit) atomicAdd(&hist[giant_part(d, I.lgp)], 1u);
        wave_append(emit, ((unsigned long long)d << 32) | c_val[i], &s_n, out);
    }
    __syncthreads();
    if (threadIdx.x == 0) h.tcnt[blockIdx.x] = s_n;
    for (uint32_t p = threadIdx.x; p < np; p += kBlock) {
        const uint32_t c = hist[p];
        if (c) h.base[I.mbase + (uint64_t)lt * np + p] = atomicAdd(&h.pcnt[I.pbase + p], c);
    }
}

__global__ __launch_bounds__(kBlock) void k_giant_scatter(GiantCtx h) {
    extern __shared__ uint32_t s_dyn[];   // [2 * h.max_parts]
    uint32_t *s_off = s_dyn;
    uint32_t *s_rank = s_dyn + h.max_parts;
    const uint64_t tile = h.tiles[blockIdx.x];
    const uint32_t lt = (uint32_t)tile;
    const GiantInfo I = h.info[tile >> 32];
    const uint32_t np = 1u << I.lgp;
    for (uint32_t p = threadIdx.x; p < np; p += kBlock) {
        // a run start exists only for partitions this tile has entries in
        s_off[p] = h.poff[I.pbase + p] + h.base[I.mbase + (uint64_t)lt * np + p];
        s_rank[p] = 0u;
    }
    __syncthreads();
    const uint32_t cnt = h.tcnt[blockIdx.x];
    const unsigned long long *in = h.p0 + I.off + ((uint64_t)lt << I.lgt);
    for (uint32_t kb = threadIdx.x; kb < cnt; kb += kBlock * kGiantIlp) {
        unsigned long long pr[kGiantIlp];
#pragma unroll
        for (int u = 0; u < kGiantIlp; ++u) {
            const uint32_t k = kb + (uint32_t)u * kBlock;
            pr[u] = k < cnt ? in[k] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kGiantIlp; ++u) {
            if (kb + (uint32_t)u * kBlock >= cnt) continue;
            const uint32_t p = giant_part((uint32_t)(pr[u] >> 32), I.lgp);
            h.p1[s_off[p] + atomicAdd(&s_rank[p], 1u)] = pr[u];
        }
    }
}

// DEC != 0: every distinct destination d != c of the partition is also scored
// as a candidate of the next level's first sweep (FirstSweep; d's new degree
// gathered), and the partition's best and non-self count are kept for
// k_giant_decide.
template <int DEC>
__global__ __launch_bounds__(kBlock) void k_giant_reduce(GiantCtx h) {
    using FS = FirstSweep<DEC == 0 ? 1 : DEC>;
    using S = typename FS::S;
    __shared__ SumTable t;
    const uint64_t item = h.items[blockIdx.x];
    const uint32_t p = (uint32_t)item;
    const GiantInfo I = h.info[item >> 32];
    sum_init(t);
    const uint32_t beg = h.poff[I.pbase + p];
    const uint32_t np = h.pcnt[I.pbase + p];
    const unsigned long long *pair = h.p1 + beg;
    unsigned long long *out = h.p0 + beg;
    const uint64_t span_lo = (uint64_t)p << (32 - I.lgp);
    const uint64_t span_hi = (uint64_t)(p + 1) << (32 - I.lgp);
    uint64_t width = sum_first_width(span_lo, span_hi, np);
    uint64_t lo = span_lo;
    uint32_t n_out = 0;
    S best = FS::SC::none();
    uint32_t best_c = LV_EMPTY;
    uint32_t nself = 0;
    const uint32_t kc = DEC ? h.a.knew[I.c] : 0u;
    int64_t m2 = h.a.m2;
    double tw = h.a.tw_total;
    const bool live = DEC && vertex_graph(h.a.dbt, I.c, h.a.m2, h.a.tw_total, m2, tw);
    __syncthreads();
    while (np > 0 && lo < span_hi) {
        const uint64_t rhi = span_hi - lo < width ? span_hi : lo + width;
        const uint32_t nu = sum_piece<kGiantIlp>(t, pair, np, lo, rhi);
        if (nu == 0xFFFFFFFFu) {   // too many destinations: halve the piece
            width = (rhi - lo) / 2 > 0 ? (rhi - lo) / 2 : 1u;
            continue;
        }
        if constexpr (DEC == 0) {
            for (uint32_t i = threadIdx.x; i < nu; i += kBlock) {
                const uint32_t s = t.used[i];
                out[n_out + i] = ((unsigned long long)t.keys[s] << 32) | t.vals[s];
                t.keys[s] = LV_EMPTY;
                t.vals[s] = 0u;
            }
        } else {
            // kGiantIlp destinations per lane, their degrees gathered together
            for (uint32_t ib = threadIdx.x; ib < nu; ib += kBlock * kGiantIlp) {
                uint32_t ds[kGiantIlp], ws[kGiantIlp], kds[kGiantIlp];
#pragma unroll
                for (int u = 0; u < kGiantIlp; ++u) {
                    const uint32_t i = ib + (uint32_t)u * kBlock;
                    ds[u] = LV_EMPTY;
                    ws[u] = 0u;
                    if (i < nu) {
                        const uint32_t s = t.used[i];
                        ds[u] = t.keys[s];
                        ws[u] = t.vals[s];
                        out[n_out + i] = ((unsigned long long)ds[u] << 32) | ws[u];
                        t.keys[s] = LV_EMPTY;
                        t.vals[s] = 0u;
                    }
                }
#pragma unroll
                for (int u = 0; u < kGiantIlp; ++u)
                    kds[u] = live && ds[u] != LV_EMPTY && ds[u] != I.c ? h.a.knew[ds[u]] : 0u;
#pragma unroll
                for (int u = 0; u < kGiantIlp; ++u) {
                    if (!live || ds[u] == LV_EMPTY || ds[u] == I.c) continue;
                    ++nself;
                    const S sc = FS::score(ws[u], kc, kds[u], m2, tw);
                    if (better_candidate(sc, ds[u], best, best_c, false)) { best = sc; best_c = ds[u]; }
                }
            }
        }
        sum_next_piece(t);
        n_out += nu;
        lo = rhi;
        if (nu < kSumTarget / 2 && width < span_hi - span_lo) width *= 2;   // sparse: widen again
    }
    if (threadIdx.x == 0) {
        h.dcnt[I.pbase + p] = n_out;
        if (n_out) atomicAdd(&h.a.rcount[I.c], n_out);
    }
    if constexpr (DEC != 0) {
        __shared__ S r_s[kBlock / 64];
        __shared__ uint32_t r_c[kBlock / 64];
        __shared__ uint32_t r_n[kBlock / 64];
        group_best<64, S>(best, best_c, false);
        const uint32_t wn = group_sum_u32<64>(nself);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; r_n[w] = wn; }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tn = 0;
            for (int i = 0; i < kBlock / 64; ++i) {
                tn += r_n[i];
                if (better_candidate(r_s[i], r_c[i], best, best_c, false)) {
                    best = r_s[i];
                    best_c = r_c[i];
                }
            }
            if constexpr (DEC == 1) h.dpart_s[I.pbase + p] = best;
            else h.dpart_s[I.pbase + p] = (int64_t)__float_as_int(best);
            h.dpart_c[I.pbase + p] = best_c;
            h.dpart_n[I.pbase + p] = tn;
        }
    }
}

// One workgroup per giant community: its partitions' bests reduced, the next
// level's first-sweep decision and statistics.
template <int DEC>
__global__ __launch_bounds__(kBlock) void k_giant_decide(GiantCtx h) {
    using FS = FirstSweep<DEC>;
    using S = typename FS::S;
    const GiantInfo I = h.info[blockIdx.x];
    const uint32_t np = 1u << I.lgp;
    S best = FS::SC::none();
    uint32_t best_c = LV_EMPTY;
    uint32_t nself = 0;
    for (uint32_t q = threadIdx.x; q < np; q += kBlock) {
        const int64_t raw = h.dpart_s[I.pbase + q];
        S sc;
        if constexpr (DEC == 1) sc = raw;
        else sc = __int_as_float((int)raw);
        const uint32_t cq = h.dpart_c[I.pbase + q];
        nself += h.dpart_n[I.pbase + q];
        if (better_candidate(sc, cq, best, best_c, false)) { best = sc; best_c = cq; }
    }
    __shared__ S r_s[kBlock / 64];
    __shared__ uint32_t r_c[kBlock / 64];
    __shared__ uint32_t r_n[kBlock / 64];
    group_best<64, S>(best, best_c, false);
    const uint32_t wn = group_sum_u32<64>(nself);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { r_s[w] = best; r_c[w] = best_c; r_n[w] = wn; }
    __syncthreads();
    uint32_t tn = 0, considered = 0, twrite = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < kBlock / 64; ++i) {
            tn += r_n[i];
            if (better_candidate(r_s[i], r_c[i], best, best_c, false)) {
                best = r_s[i];
                best_c = r_c[i];
            }
        }
        const uint32_t c = I.c;
        int64_t m2 = h.a.m2;
        double tw = h.a.tw_total;
        const bool live = vertex_graph(h.a.dbt, c, h.a.m2, h.a.tw_total, m2, tw);
        const uint32_t t = live ? FS::decide(c, h.a.knew[c], best, best_c, m2) : c;
        h.a.dtarget[c] = t;
        considered = tn > 0;
        twrite = t != c;
    }
    block_add_stats(h.a.dstats, ST_E_ACTIVE, tn, ST_CANDIDATES, tn, ST_CONSIDERED, considered,
                    ST_TWRITES, twrite);
}

__global__ __launch_bounds__(kBlock) void k_giant_copy(GiantCtx h, const uint64_t *out_row,
                                                       uint32_t *col, uint32_t *w) {
    const uint64_t item = h.items[blockIdx.x];
    const uint32_t p = (uint32_t)item;
    const GiantInfo I = h.info[item >> 32];
    const unsigned long long *in = h.p0 + h.poff[I.pbase + p];
    const uint64_t dst = out_row[I.c] + (h.doff[I.pbase + p] - h.doff[I.pbase]);
    const uint32_t n = h.dcnt[I.pbase + p];
    // kCopyIlp loads per lane before their stores: the output may alias the
    // input as far as the compiler knows, so a plain loop waits on every load
    for (uint32_t t0 = threadIdx.x; t0 < n; t0 += kBlock * kCopyIlp) {
        unsigned long long pr[kCopyIlp];
#pragma unroll
        for (int u = 0; u < kCopyIlp; ++u) {
            const uint32_t t = t0 + (uint32_t)u * kBlock;
            pr[u] = t < n ? in[t] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kCopyIlp; ++u) {
            const uint32_t t = t0 + (uint32_t)u * kBlock;
            if (t < n) {
                col[dst + t] = (uint32_t)(pr[u] >> 32);
                w[dst + t] = (uint32_t)pr[u];
            }
        }
    }
}

// Runs of equal keys in sorted keys (the sort sweep, .edge ingest): head flags.
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

// Copy the staged row of every LDS-table community to its final offset.
template <int G>
__global__ __launch_bounds__(kBlock) void k_agg_copy(AggCtx a, const uint32_t *__restrict__ list,
                                                     uint32_t count, uint32_t parts,
                                                     const uint64_t *out_row, uint32_t *col,
                                                     uint32_t *w) {
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
    const uint64_t src = a.crow[c];
    // kCopyIlp loads per lane before their stores: the output may alias the
    // staging rows as far as the compiler knows, so a plain loop waits on
    // every load
    for (uint32_t t0 = part * G + lane; t0 < r; t0 += stride * kCopyIlp) {
        uint32_t cs[kCopyIlp], ws[kCopyIlp];
#pragma unroll
        for (int u = 0; u < kCopyIlp; ++u) {
            const uint32_t t = t0 + (uint32_t)u * stride;
            cs[u] = t < r ? a.tcol[src + t] : 0u;
            ws[u] = t < r ? a.tw[src + t] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kCopyIlp; ++u) {
            const uint32_t t = t0 + (uint32_t)u * stride;
            if (t < r) {
                col[dst + t] = cs[u];
                w[dst + t] = ws[u];
            }
        }
    }
}

__global__ void k_newc(const uint32_t *comm, const uint32_t *map, uint32_t n, uint32_t *newc,
                       uint64_t *key, uint32_t *val, const uint32_t *tot,
                       unsigned long long *newck) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        const uint32_t cv = comm[v];
        const uint32_t c = renumbered(map, cv);
        newc[v] = c;
        key[v] = c;
        val[v] = (uint32_t)v;
        if (newck) newck[v] = (unsigned long long)c | ((unsigned long long)tot[cv] << 32);
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
              uint32_t *val, const uint32_t *tot, unsigned long long *newck, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_newc, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, comm, map, n,
                       newc, key, val, tot, newck);
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
    a.m2 = x.m2; a.tw_total = x.tw_total; a.newck = x.newck; a.knew = x.knew;
    a.dtarget = x.dtarget; a.dstats = x.dstats; a.dbt = x.dbt;
    return a;
}

void agg_tables(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts, hipStream_t s) {
    const AggCtx a = ctx_of(x);
    const bool wtd = x.g.wt != nullptr;
    // lanes per community in the small buckets: a quarter of the table's
    // slots, so a wave holds many communities (measured faster than half)
    for (int b = 0; b < kHubBucket; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        const uint32_t *L = lists[b];
#define LV_AGG_D(GG, TT, DD)                                                                    \
    do {                                                                                        \
        const dim3 grid((cnt + kBlock / (GG) - 1) / (kBlock / (GG)));                           \
        if (wtd) hipLaunchKernelGGL((k_agg_table<GG, TT, true, DD>), grid, dim3(kBlock), 0, s, a, L, cnt); \
        else hipLaunchKernelGGL((k_agg_table<GG, TT, false, DD>), grid, dim3(kBlock), 0, s, a, L, cnt); \
    } while (0)
#define LV_AGG(GG, TT)                                                                          \
    do {                                                                                        \
        if (x.dgain == 0) LV_AGG_D(GG, TT, 1);                                                  \
        else if (x.dgain == 1) LV_AGG_D(GG, TT, 2);                                             \
        else LV_AGG_D(GG, TT, 0);                                                               \
    } while (0)
        // tables of table_slots(the bucket's largest community) (bucket 7:
        // kTable7), the slots of the best-move tables of the same degree bound
        switch (b) {
            case 0: LV_AGG(2, kTableT[0]); break;
            case 1: LV_AGG(4, kTableT[1]); break;
            case 2: LV_AGG(8, kTableT[2]); break;
            case 3: LV_AGG(16, kTableT[3]); break;
            case 4: LV_AGG(32, kTableT[4]); break;
            case 5: LV_AGG(64, kTableT[5]); break;
            case 6: LV_AGG(256, kTableT[6]); break;
            default: LV_AGG(256, kTable7); break;
        }
#undef LV_AGG
#undef LV_AGG_D
    }
}

static GiantCtx giant_ctx(const AggArgs &x, const GiantArgs &g) {
    GiantCtx h;
    static_cast<GiantArgs &>(h) = g;
    h.a = ctx_of(x);
    return h;
}

void agg_giant(const AggArgs &x, const GiantArgs &g, uint32_t n_parts, hipStream_t s) {
    if (!g.n_tiles) return;
    const GiantCtx h = giant_ctx(x, g);
    // the per-partition LDS arrays are as large as the plan's largest community needs
    const size_t lds1 = (size_t)g.max_parts * sizeof(uint32_t);
    if (x.g.wt) hipLaunchKernelGGL(k_giant_gather<true>, dim3(g.n_tiles), dim3(kBlock), lds1, s, h);
    else hipLaunchKernelGGL(k_giant_gather<false>, dim3(g.n_tiles), dim3(kBlock), lds1, s, h);
    scan_u32(g.pcnt, g.poff, n_parts, false, g.scan_tmp, s);
    hipLaunchKernelGGL(k_giant_scatter, dim3(g.n_tiles), dim3(kBlock), 2 * lds1, s, h);
    if (x.dgain == 0) {
        hipLaunchKernelGGL(k_giant_reduce<1>, dim3(g.n_items), dim3(kBlock), 0, s, h);
        hipLaunchKernelGGL(k_giant_decide<1>, dim3(g.n_giants), dim3(kBlock), 0, s, h);
    } else if (x.dgain == 1) {
        hipLaunchKernelGGL(k_giant_reduce<2>, dim3(g.n_items), dim3(kBlock), 0, s, h);
        hipLaunchKernelGGL(k_giant_decide<2>, dim3(g.n_giants), dim3(kBlock), 0, s, h);
    } else {
        hipLaunchKernelGGL(k_giant_reduce<0>, dim3(g.n_items), dim3(kBlock), 0, s, h);
    }
    scan_u32(g.dcnt, g.doff, n_parts, false, g.scan_tmp, s);
}

void agg_giant_copy(const AggArgs &x, const GiantArgs &g, uint32_t n_parts, const uint64_t *out_row,
                    uint32_t *col, uint32_t *w, hipStream_t s) {
    (void)n_parts;
    if (!g.n_items) return;
    hipLaunchKernelGGL(k_giant_copy, dim3(g.n_items), dim3(kBlock), 0, s, giant_ctx(x, g), out_row,
                       col, w);
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

void agg_copy(const AggArgs &x, uint32_t *const *lists, const uint32_t *counts,
              const uint64_t *out_row, uint32_t *col, uint32_t *w, hipStream_t s) {
    const AggCtx a = ctx_of(x);
    for (int b = 0; b < kHubBucket; ++b) {
        const uint32_t cnt = counts[b];
        if (!cnt) continue;
        const uint32_t *L = lists[b];
#define LV_CP(GG, PARTS)                                                                        \
    do {                                                                                        \
        const uint32_t gpb = kBlock / (GG);                                                     \
        const dim3 grid((GG) == kBlock ? cnt * (PARTS) : (cnt + gpb - 1) / gpb);                \
        hipLaunchKernelGGL((k_agg_copy<GG>), grid, dim3(kBlock), 0, s, a, L, cnt,               \
                           (uint32_t)(PARTS), out_row, col, w);                                 \
    } while (0)
        switch (b) {
            case 0: LV_CP(2, 1); break;
            case 1: LV_CP(4, 1); break;
            case 2: LV_CP(8, 1); break;
            case 3: LV_CP(16, 1); break;
            case 4: LV_CP(32, 1); break;
            case 5: LV_CP(32, 1); break;
            case 6: LV_CP(256, 1); break;
            default: LV_CP(256, 4); break;
        }
#undef LV_CP
    }
}

}  // namespace lv
