// lv_server.hip -- kernels of the serving loop: the union of a level is
// assembled from segments (contracted graphs of continuing jobs, level-0
// graphs of newly admitted ones), each job is one vertex range of it, and
// every job's best_communities is composed with the level's renumbering.
#include "lv_device.h"
#include "lv_server.h"

namespace lv {

namespace {

// gid[v] = the job whose vertex range [voff[j], voff[j+1]) holds v
__global__ void k_fill_gid_ranges(const uint32_t *voff, uint32_t jobs, uint32_t n, uint32_t *gid) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = jobs;   // voff[lo] <= v < voff[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (voff[mid] <= (uint32_t)v) lo = mid; else hi = mid;
        }
        gid[v] = lo;
    }
}

// One segment per blockIdx.y: rows re-based to the destination entry offset,
// columns shifted to the destination vertex range, weights (1 when the source
// has none) and degrees copied.
__global__ __launch_bounds__(kBlock) void k_segment_copy(const Segment *segs, uint64_t *row,
                                                         uint32_t *col, uint32_t *w,
                                                         uint32_t *kdeg) {
    const Segment sg = segs[blockIdx.y];
    const uint64_t e0 = sg.row[sg.src_voff];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (size_t i = t; i < sg.n; i += stride) {
        row[sg.dst_voff + i] = sg.row[sg.src_voff + i] - e0 + sg.dst_eoff;
        kdeg[sg.dst_voff + i] = sg.kdeg[sg.src_voff + i];
    }
    const uint32_t shift = sg.dst_voff - sg.src_voff;   // modulo 2^32
    for (size_t j = t; j < sg.E; j += stride) {
        col[sg.dst_eoff + j] = sg.col[e0 + j] + shift;
        w[sg.dst_eoff + j] = sg.w ? sg.w[e0 + j] : 1u;
    }
}

// best_j[v] = map[comm[voff_j + best_j[v]]] - 1 - base_j: the local id of the
// community that v's current vertex ended the level in.
__global__ void k_update_best_jobs(const JobBest *jobs, const uint32_t *comm, const uint32_t *map) {
    const JobBest jb = jobs[blockIdx.y];
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < jb.n0;
         v += (size_t)gridDim.x * blockDim.x)
        jb.best[v] = map[comm[jb.voff + jb.best[v]]] - 1u - jb.base;
}

__global__ void k_gather_u64(const uint64_t *src, const uint32_t *idx, uint32_t n, uint64_t *out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

}  // namespace

void gather_u64(const uint64_t *src, const uint32_t *idx, uint32_t n, uint64_t *out, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, src, idx, n, out);
}

void fill_gid_ranges(const uint32_t *voff, uint32_t jobs, uint32_t n, uint32_t *gid, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_fill_gid_ranges, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s, voff,
                       jobs, n, gid);
}

void segment_copy(const Segment *segs, uint32_t count, uint64_t max_len, uint64_t *row,
                  uint32_t *col, uint32_t *w, uint32_t *kdeg, hipStream_t s) {
    const unsigned gx = grid_for(max_len ? max_len : 1, kBlock * 8, 4096u);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535u) {   // grid y is at most 65535
        const uint32_t c = count - c0 < 65535u ? count - c0 : 65535u;
        hipLaunchKernelGGL(k_segment_copy, dim3(gx, c), dim3(kBlock), 0, s, segs + c0, row, col, w,
                           kdeg);
    }
}

void update_best_jobs(const JobBest *jobs, uint32_t count, uint32_t max_n0, const uint32_t *comm,
                      const uint32_t *map, hipStream_t s) {
    const unsigned gx = grid_for(max_n0 ? max_n0 : 1, kBlock * 4, 4096u);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535u) {
        const uint32_t c = count - c0 < 65535u ? count - c0 : 65535u;
        hipLaunchKernelGGL(k_update_best_jobs, dim3(gx, c), dim3(kBlock), 0, s, jobs + c0, comm, map);
    }
}

}  // namespace lv
