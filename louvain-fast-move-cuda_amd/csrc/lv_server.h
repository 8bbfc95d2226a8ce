// lv_server.h -- device records and launchers of the serving loop (lv_server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lv {

// One piece of the next union: vertices [src_voff, src_voff + n) of a source
// CSR (their entries are contiguous) go to [dst_voff, dst_voff + n).
struct Segment {
    const uint64_t *row;
    const uint32_t *col;
    const uint32_t *w;        // nullptr: unit weights
    const uint32_t *kdeg;     // nullptr: degrees are computed afterwards
    uint32_t src_voff, dst_voff, n;
    uint64_t dst_eoff, E;
};

// One job's best_communities and its place in the level that just ended.
struct JobBest {
    uint32_t *best;           // [n0] local ids
    uint32_t n0;
    uint32_t voff;            // first union vertex of the job at this level
    uint32_t base;            // renumbered ids of communities before the job's
};

// A just-admitted job's level-0 CSR, uploaded raw (job-local offsets and ids)
// into the union at [dst_voff, dst_voff + n) / [dst_eoff, dst_eoff + E).
struct FreshSeg {
    uint32_t dst_voff, n, slot;
    uint32_t unit;            // 1: no weights were uploaded (all 1)
    uint64_t dst_eoff, E;
};

// One job's best_communities moved to the new arena (or set to the identity).
struct BestMove {
    const uint32_t *src;      // nullptr: identity
    uint32_t *dst;
    uint32_t n;
};

// Shift the raw rows and columns of fresh segments into union coordinates
// (and write unit weights).
void fresh_shift(const FreshSeg *fs, uint32_t count, uint64_t max_len, uint64_t *row,
                 uint32_t *col, uint32_t *w, hipStream_t s);
// Degrees of fresh vertices, and per job the identity partition's parts:
// sig[slot] += self-loop weight, sq[slot] += k^2 (rows must be final).
void fresh_parts(const FreshSeg *fs, uint32_t count, uint32_t max_n, uint64_t max_e,
                 const uint64_t *row, const uint32_t *col, const uint32_t *w, uint32_t *kdeg,
                 unsigned long long *sig, unsigned long long *sq, hipStream_t s);
void move_best(const BestMove *mv, uint32_t count, uint32_t max_n, hipStream_t s);
// out[i] = src[idx[i]]
void gather_u64(const uint64_t *src, const uint32_t *idx, uint32_t n, uint64_t *out, hipStream_t s);
// gid[v] = j with voff[j] <= v < voff[j+1] (voff has jobs+1 entries)
void fill_gid_ranges(const uint32_t *voff, uint32_t jobs, uint32_t n, uint32_t *gid, hipStream_t s);
// segs[count] in device memory; max_len = largest n or E of a segment
void segment_copy(const Segment *segs, uint32_t count, uint64_t max_len, uint64_t *row,
                  uint32_t *col, uint32_t *w, uint32_t *kdeg, hipStream_t s);
// jobs[count] in device memory; max_n0 = largest n0
void update_best_jobs(const JobBest *jobs, uint32_t count, uint32_t max_n0, const uint32_t *comm,
                      const uint32_t *map, hipStream_t s);

}  // namespace lv
