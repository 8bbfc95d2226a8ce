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

constexpr int kSegIlp = 4;   // entries per lane with loads in flight together

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
        if (sg.kdeg) kdeg[sg.dst_voff + i] = sg.kdeg[sg.src_voff + i];
    }
    const uint32_t shift = sg.dst_voff - sg.src_voff;   // modulo 2^32
    // kSegIlp loads per lane before their stores (the compiler cannot move a
    // load above a store that may alias it)
    for (size_t j0 = t; j0 < sg.E; j0 += stride * kSegIlp) {
        uint32_t cs[kSegIlp], ws[kSegIlp];
#pragma unroll
        for (int u = 0; u < kSegIlp; ++u) {
            const size_t j = j0 + (size_t)u * stride;
            cs[u] = j < sg.E ? sg.col[e0 + j] : 0u;
            ws[u] = (j < sg.E && sg.w) ? sg.w[e0 + j] : 1u;
        }
#pragma unroll
        for (int u = 0; u < kSegIlp; ++u) {
            const size_t j = j0 + (size_t)u * stride;
            if (j < sg.E) {
                col[sg.dst_eoff + j] = cs[u] + shift;
                w[sg.dst_eoff + j] = ws[u];
            }
        }
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

__global__ __launch_bounds__(kBlock) void k_fresh_shift(const FreshSeg *segs, uint64_t *row,
                                                        uint32_t *col, uint32_t *w) {
    const FreshSeg sg = segs[blockIdx.y];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (size_t i = t; i < sg.n; i += stride) row[sg.dst_voff + i] += sg.dst_eoff;
    for (size_t j0 = t; j0 < sg.E; j0 += stride * kSegIlp) {
        uint32_t cs[kSegIlp];
#pragma unroll
        for (int u = 0; u < kSegIlp; ++u) {
            const size_t j = j0 + (size_t)u * stride;
            cs[u] = j < sg.E ? col[sg.dst_eoff + j] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kSegIlp; ++u) {
            const size_t j = j0 + (size_t)u * stride;
            if (j < sg.E) {
                col[sg.dst_eoff + j] = cs[u] + sg.dst_voff;
                if (sg.unit) w[sg.dst_eoff + j] = 1u;
            }
        }
    }
}

// Degrees of fresh vertices, entry-parallel: each thread sums kChunkE entries
// of one segment, flushing a vertex's partial sum with one atomic (kdeg of
// the segment's vertices is zero beforehand); self-loop weights go to sig.
constexpr int kChunkE = 16;

__global__ __launch_bounds__(kBlock) void k_fresh_degrees(const FreshSeg *segs, const uint64_t *row,
                                                          const uint32_t *col, const uint32_t *w,
                                                          uint32_t *kdeg, unsigned long long *sig) {
    const FreshSeg sg = segs[blockIdx.y];
    uint64_t s_self = 0;
    const uint64_t nch = (sg.E + kChunkE - 1) / kChunkE;
    for (uint64_t ch = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; ch < nch;
         ch += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j0 = sg.dst_eoff + ch * kChunkE;
        const uint64_t j1 = j0 + kChunkE < sg.dst_eoff + sg.E ? j0 + kChunkE : sg.dst_eoff + sg.E;
        // vertex whose row holds j0: last v in the segment with row[v] <= j0
        uint32_t lo = sg.dst_voff, hi = sg.dst_voff + sg.n;
        while (hi - lo > 1) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (row[mid] <= j0) lo = mid; else hi = mid;
        }
        uint32_t v = lo;
        uint64_t vend = row[v + 1];
        uint32_t acc = 0;
        for (uint64_t j = j0; j < j1; ++j) {
            while (j >= vend) {
                if (acc) atomicAdd(&kdeg[v], acc);
                acc = 0;
                ++v;
                vend = row[v + 1];
            }
            const uint32_t x = w[j];
            acc += x;
            if (col[j] == v) s_self += x;
        }
        if (acc) atomicAdd(&kdeg[v], acc);
    }
    s_self = group_sum_u64<64>(s_self);
    if ((threadIdx.x & 63) == 0 && s_self) atomicAdd(&sig[sg.slot], (unsigned long long)s_self);
}

__global__ __launch_bounds__(kBlock) void k_fresh_sumsq(const FreshSeg *segs, const uint32_t *kdeg,
                                                        unsigned long long *sq) {
    const FreshSeg sg = segs[blockIdx.y];
    uint64_t s_sq = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < sg.n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = kdeg[sg.dst_voff + i];
        s_sq += k * k;
    }
    s_sq = group_sum_u64<64>(s_sq);
    if ((threadIdx.x & 63) == 0 && s_sq) atomicAdd(&sq[sg.slot], (unsigned long long)s_sq);
}

__global__ void k_fresh_zero(const FreshSeg *segs, uint32_t *kdeg) {
    const FreshSeg sg = segs[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < sg.n;
         i += (size_t)gridDim.x * blockDim.x)
        kdeg[sg.dst_voff + i] = 0u;
}

__global__ void k_move_best(const BestMove *mv) {
    const BestMove m = mv[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m.n;
         i += (size_t)gridDim.x * blockDim.x)
        m.dst[i] = m.src ? m.src[i] : (uint32_t)i;
}

}  // namespace

void fresh_shift(const FreshSeg *fs, uint32_t count, uint64_t max_len, uint64_t *row,
                 uint32_t *col, uint32_t *w, hipStream_t s) {
    const unsigned gx = grid_for(max_len ? max_len : 1, kBlock * 8, 4096u);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535u) {
        const uint32_t c = count - c0 < 65535u ? count - c0 : 65535u;
        hipLaunchKernelGGL(k_fresh_shift, dim3(gx, c), dim3(kBlock), 0, s, fs + c0, row, col, w);
    }
}

void fresh_parts(const FreshSeg *fs, uint32_t count, uint32_t max_n, uint64_t max_e,
                 const uint64_t *row, const uint32_t *col, const uint32_t *w, uint32_t *kdeg,
                 unsigned long long *sig, unsigned long long *sq, hipStream_t s) {
    const unsigned gn = grid_for(max_n ? max_n : 1, kBlock * 4, 4096u);
    const unsigned ge = grid_for(max_e ? (max_e + kChunkE - 1) / kChunkE : 1, kBlock, 4096u);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535u) {
        const uint32_t c = count - c0 < 65535u ? count - c0 : 65535u;
        hipLaunchKernelGGL(k_fresh_zero, dim3(gn, c), dim3(kBlock), 0, s, fs + c0, kdeg);
        hipLaunchKernelGGL(k_fresh_degrees, dim3(ge, c), dim3(kBlock), 0, s, fs + c0, row, col, w,
                           kdeg, sig);
        hipLaunchKernelGGL(k_fresh_sumsq, dim3(gn, c), dim3(kBlock), 0, s, fs + c0, kdeg, sq);
    }
}

void move_best(const BestMove *mv, uint32_t count, uint32_t max_n, hipStream_t s) {
    const unsigned gx = grid_for(max_n ? max_n : 1, kBlock * 4, 4096u);
    for (uint32_t c0 = 0; c0 < count; c0 += 65535u) {
        const uint32_t c = count - c0 < 65535u ? count - c0 : 65535u;
        hipLaunchKernelGGL(k_move_best, dim3(gx, c), dim3(kBlock), 0, s, mv + c0);
    }
}

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
