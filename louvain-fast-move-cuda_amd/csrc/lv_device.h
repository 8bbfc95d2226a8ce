// lv_device.h -- device helpers shared by the gfx950 kernels: wave64
// shuffles and reductions, LDS hash slots, per-block statistics, degree
// buckets and bucket-list appends.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "../../include/constants.h"
#include "../../include/operators.h"
#include "lv_kernels.h"

namespace lv {

constexpr int kBlock = 256;

__device__ inline unsigned long long lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : ((~0ull) >> (64 - lane));
}

__device__ inline int64_t shfl_xor_i64(int64_t v, int off) {
    const int lo = __shfl_xor((int)(uint32_t)(uint64_t)v, off);
    const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), off);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int off) {
    return (uint64_t)shfl_xor_i64((int64_t)v, off);
}

__device__ inline float shfl_xor_f32(float v, int off) { return __shfl_xor(v, off); }

template <typename S>
__device__ inline S shfl_xor_s(S v, int off);
template <>
__device__ inline int64_t shfl_xor_s<int64_t>(int64_t v, int off) { return shfl_xor_i64(v, off); }
template <>
__device__ inline float shfl_xor_s<float>(float v, int off) { return shfl_xor_f32(v, off); }

// Index of a lane's group of G lanes in its workgroup.  A group of a whole
// wave or more is wave uniform, which the compiler is told (readfirstlane):
// the group's words then load through the scalar unit.
template <int G>
__device__ inline uint32_t group_index() {
    const uint32_t g = threadIdx.x / (uint32_t)G;
    if constexpr (G >= 64) return (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
    else return g;
}

// ---- reductions inside aligned groups of G <= 64 lanes ----
//
// Inside a row of 16 lanes the partner of a butterfly step comes by DPP (a
// VALU operand modifier, no LDS round trip): quad permutations for the
// partners at distance 1 and 2, then the half-row and row mirrors (lane i with
// 7 - i, then with 15 - i), which pair every lane with one of the other half
// as a butterfly must.  After the four steps every lane of a row holds the
// row's result; rows are combined through scalar reads of lanes 0, 16, 32, 48.
// The operations reduced are commutative and associative (sums, the best of a
// total order), so the pairing order does not change the result.  Every lane
// of the wave must be active.
constexpr int kDppXor1 = 0xB1;          // quad_perm(1, 0, 3, 2)
constexpr int kDppXor2 = 0x4E;          // quad_perm(2, 3, 0, 1)
constexpr int kDppHalfMirror = 0x141;   // row_half_mirror
constexpr int kDppMirror = 0x140;       // row_mirror

template <int CTRL>
__device__ inline uint32_t dpp_u32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ inline uint64_t dpp_u64(uint64_t x) {
    return (uint64_t)dpp_u32<CTRL>((uint32_t)x) |
           ((uint64_t)dpp_u32<CTRL>((uint32_t)(x >> 32)) << 32);
}

__device__ inline uint32_t lane_u32(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

__device__ inline uint64_t lane_u64(uint64_t x, int lane) {
    return (uint64_t)lane_u32((uint32_t)x, lane) | ((uint64_t)lane_u32((uint32_t)(x >> 32), lane) << 32);
}

// score words: int64 (exact gains) or f32 (the reference's rounding)
template <typename S>
__device__ inline uint64_t s_bits(S s) {
    if constexpr (sizeof(S) == 8) return (uint64_t)s;
    else return (uint64_t)__float_as_uint(s);
}
template <typename S>
__device__ inline S s_from(uint64_t b) {
    if constexpr (sizeof(S) == 8) return (S)b;
    else return __uint_as_float((uint32_t)b);
}

template <int CTRL, typename S>
__device__ inline void best_step(S &s, uint32_t &c, bool highest) {
    S os;
    if constexpr (sizeof(S) == 8) os = s_from<S>(dpp_u64<CTRL>(s_bits(s)));
    else os = s_from<S>(dpp_u32<CTRL>((uint32_t)s_bits(s)));
    const uint32_t oc = dpp_u32<CTRL>(c);
    if (better_candidate(os, oc, s, c, highest)) { s = os; c = oc; }
}

template <typename S>
__device__ inline void best_of_lane(S &s, uint32_t &c, const S &s0, uint32_t c0, int lane,
                                    bool highest) {
    S os;
    if constexpr (sizeof(S) == 8) os = s_from<S>(lane_u64(s_bits(s0), lane));
    else os = s_from<S>(lane_u32((uint32_t)s_bits(s0), lane));
    const uint32_t oc = lane_u32(c0, lane);
    if (better_candidate(os, oc, s, c, highest)) { s = os; c = oc; }
}

// (score, community) best of each aligned group of G lanes (G <= 64); every
// lane of the group ends with the group's best.
template <int G, typename S>
__device__ inline void group_best(S &s, uint32_t &c, bool highest) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "groups of 1-64 lanes");
    if constexpr (G >= 2) best_step<kDppXor1>(s, c, highest);
    if constexpr (G >= 4) best_step<kDppXor2>(s, c, highest);
    if constexpr (G >= 8) best_step<kDppHalfMirror>(s, c, highest);
    if constexpr (G >= 16) best_step<kDppMirror>(s, c, highest);
    if constexpr (G >= 32) {
        const S rs = s;
        const uint32_t rc = c;
        const bool up = (threadIdx.x & 32) != 0;
        // G == 32: rows 0-1 for the lower half, rows 2-3 for the upper
        S a = s_from<S>(0), b = s_from<S>(0);
        uint32_t ac = 0, bc = 0;
        if constexpr (sizeof(S) == 8) {
            a = s_from<S>(lane_u64(s_bits(rs), 0)); b = s_from<S>(lane_u64(s_bits(rs), 32));
        } else {
            a = s_from<S>(lane_u32((uint32_t)s_bits(rs), 0));
            b = s_from<S>(lane_u32((uint32_t)s_bits(rs), 32));
        }
        ac = lane_u32(rc, 0);
        bc = lane_u32(rc, 32);
        best_of_lane(a, ac, rs, rc, 16, highest);
        best_of_lane(b, bc, rs, rc, 48, highest);
        if constexpr (G == 64) {
            if (better_candidate(b, bc, a, ac, highest)) { a = b; ac = bc; }
            s = a;
            c = ac;
        } else {
            s = up ? b : a;
            c = up ? bc : ac;
        }
    }
}

// Sum of each aligned group of G lanes (G <= 64), in every lane of the group.
template <int G>
__device__ inline uint64_t group_sum_u64(uint64_t x) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "groups of 1-64 lanes");
    if constexpr (G >= 2) x += dpp_u64<kDppXor1>(x);
    if constexpr (G >= 4) x += dpp_u64<kDppXor2>(x);
    if constexpr (G >= 8) x += dpp_u64<kDppHalfMirror>(x);
    if constexpr (G >= 16) x += dpp_u64<kDppMirror>(x);
    if constexpr (G >= 32) {
        const uint64_t lo = lane_u64(x, 0) + lane_u64(x, 16);
        const uint64_t hi = lane_u64(x, 32) + lane_u64(x, 48);
        if constexpr (G == 64) x = lo + hi;
        else x = (threadIdx.x & 32) ? hi : lo;
    }
    return x;
}

// Sum of each aligned group of G lanes (G <= 64) of 32-bit values, in every
// lane of the group; for counts whose group total stays below 2^32.
template <int G>
__device__ inline uint32_t group_sum_u32(uint32_t x) {
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "groups of 1-64 lanes");
    if constexpr (G >= 2) x += dpp_u32<kDppXor1>(x);
    if constexpr (G >= 4) x += dpp_u32<kDppXor2>(x);
    if constexpr (G >= 8) x += dpp_u32<kDppHalfMirror>(x);
    if constexpr (G >= 16) x += dpp_u32<kDppMirror>(x);
    if constexpr (G >= 32) {
        const uint32_t lo = lane_u32(x, 0) + lane_u32(x, 16);
        const uint32_t hi = lane_u32(x, 32) + lane_u32(x, 48);
        if constexpr (G == 64) x = lo + hi;
        else x = (threadIdx.x & 32) ? hi : lo;
    }
    return x;
}

template <typename U>
__device__ inline U wave_total(U x) {
    if constexpr (sizeof(U) == 4) return group_sum_u32<64>(x);
    else return group_sum_u64<64>(x);
}

template <typename U>
struct same_type { using type = U; };

// Block-wide sum of up to 4 counters; thread 0 adds them to the spread slots
// (a negative stat index skips a counter).  32-bit counters (U = uint32_t,
// wave totals below 2^32) take half the cross-lane steps of 64-bit ones.
template <typename U>
__device__ inline void block_add_stats(unsigned long long *stats, int st0, U a0, int st1,
                                       typename same_type<U>::type a1, int st2,
                                       typename same_type<U>::type a2, int st3 = -1,
                                       typename same_type<U>::type a3 = 0) {
    __shared__ uint64_t red[4][kBlock / 64];
    const int st[4] = {st0, st1, st2, st3};
    const U a[4] = {a0, a1, a2, a3};
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (st[k] < 0) continue;
        const U t = wave_total(a[k]);
        if ((threadIdx.x & 63) == 0) red[k][w] = (uint64_t)t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int slot = blockIdx.x & (kStatSlots - 1);
        for (int k = 0; k < 4; ++k) {
            if (st[k] < 0) continue;
            uint64_t t = 0;
            for (int i = 0; i < kBlock / 64; ++i) t += red[k][i];
            if (t) atomicAdd(&stats[st[k] * kStatSlots + slot], (unsigned long long)t);
        }
    }
    __syncthreads();
}

// Whether any lane of each aligned group of G lanes (G <= 64) has `x` set:
// one ballot, no cross-lane arithmetic.
template <int G>
__device__ inline bool group_any(bool x) {
    const unsigned long long b = __ballot(x);
    if constexpr (G == 64) {
        return b != 0;
    } else {
        const uint32_t base = (threadIdx.x & 63) & ~(uint32_t)(G - 1);
        return ((b >> base) & ((1ull << G) - 1)) != 0;
    }
}

// Appends `pair` (lanes with `emit`) at out[*n ...] (n in LDS): one LDS atomic
// per wave.  Lanes of a wave that are active call it together.
__device__ inline void wave_append(bool emit, unsigned long long pair, uint32_t *n,
                                   unsigned long long *out) {
    const unsigned long long m = __ballot(emit);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(n, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, leader);
    if (emit) out[base + (uint32_t)__popcll(m & lanemask_lt())] = pair;
}

// Adds val to base[key] for the lanes with `has`.  A wave whose contributing
// lanes share one key (the usual case: lists and tiles follow vertex order and
// a graph of a batch is a vertex range) issues one atomic.  Every lane of the
// wave must call it.
__device__ inline void wave_keyed_add(unsigned long long *base, uint32_t key, uint64_t val,
                                      bool has) {
    const unsigned long long m = __ballot(has);
    if (m == 0) return;
    const int leader = __ffsll((long long)m) - 1;
    const uint32_t k0 = (uint32_t)__shfl((int)key, leader);
    if (__ballot(has && key != k0) == 0) {
        const uint64_t t = group_sum_u64<64>(has ? val : 0ull);
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&base[k0], (unsigned long long)t);
    } else if (has) {
        atomicAdd(&base[key], (unsigned long long)val);
    }
}

// 2m and m of vertex v's graph; false if that graph no longer sweeps.
__device__ inline bool vertex_graph(const BatchView &b, uint32_t v, int64_t m2, double tw,
                                    int64_t &m2v, double &twv) {
    if (!b.gid) { m2v = m2; twv = tw; return true; }
    const uint32_t g = b.gid[v];
    m2v = b.m2[g];
    twv = b.tw[g];
    return b.live[g] != 0;
}

// Adjacency streams are read once per pass: non-temporal loads keep them from
// evicting the gathered comm/tot lines from L2.
__device__ inline uint32_t stream_load(const uint32_t *p) { return __builtin_nontemporal_load(p); }

// log2 of the smallest power of two >= 2*d (at least 2 slots): an
// open-addressing table for d keys then stays at most half full.
__device__ inline uint32_t table_log2(uint64_t d) {
    const uint64_t need = 2 * d;
    return need <= 2 ? 1u : (uint32_t)(64 - __builtin_clzll(need - 1));
}

__device__ inline uint32_t slot_hash_dyn(uint32_t c, uint32_t log2t) {
    return (c * 2654435761u) >> (32 - log2t);
}

// Barrier over the G lanes that share one vertex or community: a wavefront
// barrier with LDS ordering for groups inside one wave (G <= 64), the
// workgroup barrier otherwise.
template <int G>
__device__ inline void group_barrier() {
    static_assert(G <= 64 || G == kBlock, "a group is inside one wave or the whole workgroup");
    if constexpr (G == kBlock) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Slot in [0, t) of key c for tables of any size t (multiplicative hash,
// scaled by a high multiply).
__device__ inline uint32_t slot_hash_mod(uint32_t c, uint32_t t) {
    return (uint32_t)(((uint64_t)(c * 2654435761u) * t) >> 32);
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x >> 1); }

// ---- LDS sums of (key << 32 | weight) pairs, range by range of a key mix ----
//
// Rows too long for one LDS table (hub rows of the best-move pass, giant
// communities of the contraction) are cut into partitions by the top bits of
// hub_mix(key); a workgroup then sums one partition's pairs per key in an LDS
// table.  A partition may hold more keys than the table: its mix range is then
// done in pieces, each piece inserted, scanned and cleared in turn, and a
// piece that overflows is halved and redone.  hub_mix is a bijection, so a
// piece of one mix value holds one key and the halving ends.

constexpr int kSumLds = 4096;          // table slots
constexpr uint32_t kSumFill = 3072;    // keys held before a piece counts as overflowing
constexpr uint32_t kSumTarget = 2048;  // expected keys per piece

// murmur3's 32-bit finaliser (a bijection of the 32-bit keys).
__device__ inline uint32_t hub_mix(uint32_t c) {
    c ^= c >> 16;
    c *= 0x85ebca6bu;
    c ^= c >> 13;
    c *= 0xc2b2ae35u;
    c ^= c >> 16;
    return c;
}

struct SumTable {
    uint32_t keys[kSumLds];    // LV_EMPTY when free
    uint32_t vals[kSumLds];
    uint16_t used[kSumFill];   // claimed slots in claim order
    uint32_t n_used;
};

// Workgroup barrier that orders LDS only: global loads and stores in flight
// (prefetches, output writes) stay in flight across it, where
// __syncthreads() would wait for them.  For loops whose workgroup shares
// nothing but LDS (the persistent best-move tables, the LDS sum tables).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Workgroup-wide: empty table (call once; a caller leaves it empty again by
// clearing the slots it read after each sum_piece, and an overflowing
// sum_piece clears it itself).
__device__ inline void sum_init(SumTable &t) {
    for (int i = threadIdx.x; i < kSumLds; i += blockDim.x) { t.keys[i] = LV_EMPTY; t.vals[i] = 0u; }
    if (threadIdx.x == 0) t.n_used = 0;
}

// Workgroup-wide: adds the pairs whose key mixes into [lo, hi) to the table.
// Returns the number of distinct keys (uniform), or 0xFFFFFFFF after an
// overflow, in which case the table is empty again.
template <int ILP>
__device__ inline uint32_t sum_piece(SumTable &t, const unsigned long long *__restrict__ pair,
                                     uint32_t np, uint64_t lo, uint64_t hi) {
    for (uint32_t ib = threadIdx.x; ib < np; ib += blockDim.x * ILP) {
        // an overflowing piece is redone: stop reading it
        if (__hip_atomic_load(&t.n_used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= kSumFill)
            break;
        unsigned long long pr[ILP];
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            const uint32_t i = ib + (uint32_t)u * blockDim.x;
            pr[u] = i < np ? pair[i] : ~0ull;
        }
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
            const uint32_t c = (uint32_t)(pr[u] >> 32);
            if (c == LV_EMPTY) continue;
            const uint32_t h = hub_mix(c);
            if (h < lo || h >= hi) continue;
            const uint32_t w = (uint32_t)pr[u];
            // at most one new key per lane once the piece is full, so the
            // table never fills and every probe ends
            if (__hip_atomic_load(&t.n_used, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= kSumFill)
                continue;
            uint32_t s = h & (kSumLds - 1);
            for (;;) {
                const uint32_t prev = atomicCAS(&t.keys[s], LV_EMPTY, c);
                if (prev == LV_EMPTY) {
                    atomicAdd(&t.vals[s], w);
                    const uint32_t k = atomicAdd(&t.n_used, 1u);
                    if (k < kSumFill) t.used[k] = (uint16_t)s;
                    break;
                }
                if (prev == c) { atomicAdd(&t.vals[s], w); break; }
                s = (s + 1) & (kSumLds - 1);
            }
        }
    }
    lds_barrier();
    const uint32_t nu = t.n_used;
    if (nu < kSumFill) return nu;
    for (int i = threadIdx.x; i < kSumLds; i += blockDim.x) { t.keys[i] = LV_EMPTY; t.vals[i] = 0u; }
    lds_barrier();
    if (threadIdx.x == 0) t.n_used = 0;
    lds_barrier();
    return 0xFFFFFFFFu;
}

// Workgroup-wide, after the caller has read the piece's nu used slots (and
// cleared them): resets the claim count for the next piece.
__device__ inline void sum_next_piece(SumTable &t) {
    lds_barrier();
    if (threadIdx.x == 0) t.n_used = 0;
    lds_barrier();
}

// Width of the first piece of [span_lo, span_hi) for np pairs: as if the
// pairs were distinct, but at most 4 pieces, since many pairs may share a key
// (a piece that overflows is halved, one that stays sparse lets the next grow).

__device__ inline uint64_t sum_first_width(uint64_t span_lo, uint64_t span_hi, uint32_t np) {
    uint64_t ranges = 1;
    while (ranges < 4 && ranges < span_hi - span_lo && ranges * kSumTarget < np) ranges <<= 1;
    return (span_hi - span_lo) / ranges;
}

template <int GAIN>
struct Score;
template <>
struct Score<0> {
    using T = int64_t;
    __device__ static T none() { return LLONG_MIN; }
    __device__ static T of(uint32_t w, uint32_t self, uint32_t k, uint32_t tot_own, uint32_t tot_c,
                           int64_t m2, double) {
        (void)self; (void)tot_own;
        return move_score_exact(w, k, tot_c, m2);
    }
    __device__ static bool positive(T best, uint32_t self, uint32_t k, uint32_t tot_own, int64_t m2) {
        return delta_modularity_exact(best, self, k, tot_own, m2) > 0;
    }
};
template <>
struct Score<1> {
    using T = float;
    __device__ static T none() { return -INFINITY; }
    __device__ static T of(uint32_t w, uint32_t self, uint32_t k, uint32_t tot_own, uint32_t tot_c,
                           int64_t, double tw) {
        return delta_modularity_ref_f32(w, self, k, tot_own, tot_c, tw);
    }
    __device__ static bool positive(T best, uint32_t, uint32_t, uint32_t, int64_t) {
        return best > 0.0f;
    }
};


inline unsigned grid_for(size_t n, size_t per_block, unsigned cap = 65536u) {
    size_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

__device__ inline uint8_t bucket_of_degree(uint64_t d, const BucketBounds &bounds) {
#pragma unroll
    for (int b = 0; b < kNumBuckets - 1; ++b)
        if (d <= bounds.max_deg[b]) return (uint8_t)b;
    return (uint8_t)kHubBucket;
}

struct ListPtrs {
    uint32_t *p[kNumBuckets];
};

}  // namespace lv
