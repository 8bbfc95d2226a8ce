// lv_noalloc.hip -- lv_modularity_parts_noalloc: the exact modularity parts of
// a partition on a caller's device-resident CSR (what Community::
// compute_modularity, community.h:68-106, computes for a Community), for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).
//
// Scratch: the community totals (one u32 per vertex, at most kNoAllocMaxN),
// the statistic counters of the modularity kernel and an error word live in
// the file-scope device array `scratch`.  The host never passes the array's
// name to a kernel or a copy: at the first call on a device the array's
// device address is asked of the runtime once (hipGetSymbolAddress, with that
// device current) and kept per device.  Every call clears what must start at
// zero with a fill kernel (no memset node, so the sequence can be captured).
// All calls on one device share its array: calls that overlap on one device
// (two host threads, or two streams whose work interleaves) race.
#include "../../include/lv_api.h"

#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/device_array.h"
#include "lv_device.h"
#include "lv_kernels.h"

namespace lv {

namespace {

constexpr uint32_t kNoAllocMaxN = 1u << 24;                   // 16,777,216 vertices
constexpr size_t kStatWords = (size_t)kNumStats * kStatSlots;  // modularity counters
constexpr size_t kErrWord = kStatWords;                        // community id >= n seen
constexpr size_t kTotWord = kStatWords + 1;                    // tot[] as u32 from here
constexpr size_t kScratchWords = kTotWord + kNoAllocMaxN / 2;  // 64 MiB + 3.5 KiB

__device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long *g_scratch[kMaxDevices] = {};

__global__ void k_zero_words(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// tot[comm[v]] += k_v; a community id outside [0, n) touches nothing and
// raises the error word instead.
__global__ void k_totals_checked(const uint32_t *comm, const uint32_t *kdeg, uint32_t n,
                                 uint32_t *tot, unsigned long long *err) {
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n;
         v += (size_t)gridDim.x * blockDim.x) {
        const uint32_t c = comm[v];
        if (c < n) atomicAdd(&tot[c], kdeg[v]);
        else atomicOr(err, 1ull);
    }
}

// Device address of the current device's scratch array.
unsigned long long *scratch_of_current_device() {
    int dev = 0;
    LV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(LV_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        LV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned long long *>(p);
    }
    return g_scratch[dev];
}

}  // namespace

}  // namespace lv

extern "C" lv_status lv_modularity_parts_noalloc(uint32_t n, uint64_t E, const uint64_t *row,
                                                 const uint32_t *col, const uint32_t *w,
                                                 const uint32_t *kdeg, const uint32_t *comm,
                                                 double total_weight, void *stream,
                                                 int64_t *sigma_in, uint64_t *sum_tot_sq,
                                                 double *q) {
    using namespace lv;
    if (n > kNoAllocMaxN) return LV_ERR_RANGE;
    if (n == 0 || !row || (E > 0 && !col) || !kdeg || !comm || !(total_weight > 0))
        return LV_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return LV_ERR_NO_DEVICE;
    }
    try {
        hipStream_t s = static_cast<hipStream_t>(stream);
        unsigned long long *base = scratch_of_current_device();
        unsigned long long *stats = base;
        uint32_t *tot = reinterpret_cast<uint32_t *>(base + kTotWord);
        const size_t words = kTotWord + ((size_t)n + 1) / 2;
        hipLaunchKernelGGL(k_zero_words, dim3(grid_for(words, kBlock * 4)), dim3(kBlock), 0, s,
                           base, words);
        hipLaunchKernelGGL(k_totals_checked, dim3(grid_for(n, kBlock * 4)), dim3(kBlock), 0, s,
                           comm, kdeg, n, tot, base + kErrWord);
        LevelView g;
        g.row = row;
        g.col = col;
        g.wt = w;
        g.kdeg = kdeg;
        g.n = n;
        g.E = E;
        modularity_parts(g, comm, tot, stats, s);
        LV_HIP(hipGetLastError());
        unsigned long long host[kStatWords + 1];
        LV_HIP(hipMemcpyAsync(host, base, sizeof(host), hipMemcpyDeviceToHost, s));
        LV_HIP(hipStreamSynchronize(s));
        if (host[kErrWord]) return LV_ERR_ARG;
        uint64_t sig = 0, sq = 0;
        for (int i = 0; i < kStatSlots; ++i) {
            sig += host[ST_DSIGMA * kStatSlots + i];
            sq += host[ST_SUMSQ * kStatSlots + i];
        }
        const double tw = total_weight;
        if (sigma_in) *sigma_in = (int64_t)sig;
        if (sum_tot_sq) *sum_tot_sq = sq;
        if (q) *q = (double)(int64_t)sig / (2 * tw) - (double)sq / (4 * tw * tw);  // community.h:104
        return LV_OK;
    } catch (const Error &e) {
        return e.status;
    } catch (...) {
        return LV_ERR_HIP;
    }
}
