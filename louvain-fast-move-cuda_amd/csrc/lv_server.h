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
    const uint32_t *kdeg;
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
