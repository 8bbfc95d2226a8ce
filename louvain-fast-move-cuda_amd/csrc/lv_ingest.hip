// lv_ingest.hip -- .edge text to a device CSR on the GPU (reference:
// GraphHost, graph.cu:17-87, and GraphDevice's sort, graph.cuh:46-99).
//
// The file's bytes are uploaded in double-buffered pinned chunks.  Tokens are
// the reads of `std::istream >> int` (graph.cu:36-44): a read starts at the
// first byte that is not white space, takes an optional sign and the digits
// after it and stops before the next non-digit; the next read starts there.
// Byte i starts a read iff it is not white space and i is the first byte, or
// byte i-1 is white space, or byte i-1 is a digit and byte i is not.  A read
// fails on a byte that is neither sign nor digit, on a sign not followed by a
// digit, and on a value outside int; parsing stops at the first failure, as
// the reference's loop does, and only whole pairs (triples with a weight
// column; a negative weight also stops) are kept.  Then, as GraphHost does:
// (min, max) normalisation, both directions, duplicates dropped (the first
// occurrence's weight wins: a stable sort keeps record order), a self-loop
// entered twice, ids shifted by the minimum and required to be dense.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/graph.h"
#include "lv_device.h"
#include "lv_ingest.h"

namespace lv {

namespace {

constexpr int kBytesPerThread = 64;

__device__ inline bool is_ws(uint8_t c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}
__device__ inline bool is_digit(uint8_t c) { return c >= '0' && c <= '9'; }
__device__ inline bool is_sign(uint8_t c) { return c == '+' || c == '-'; }

__device__ inline bool read_start(const uint8_t *b, uint64_t i) {
    const uint8_t c = b[i];
    if (is_ws(c)) return false;
    if (i == 0) return true;
    const uint8_t p = b[i - 1];
    return is_ws(p) || (is_digit(p) && !is_digit(c));
}

// Value of the read starting at i; false if that read fails.
__device__ inline bool read_value(const uint8_t *b, uint64_t n, uint64_t i, int32_t &out) {
    bool neg = false;
    if (is_sign(b[i])) {
        neg = b[i] == '-';
        ++i;
    }
    if (i >= n || !is_digit(b[i])) return false;
    int64_t v = 0;
    while (i < n && is_digit(b[i])) {
        v = v * 10 + (b[i] - '0');
        if (v > 2147483648ll) return false;
        ++i;
    }
    v = neg ? -v : v;
    if (v > 2147483647ll || v < -2147483648ll) return false;
    out = (int32_t)v;
    return true;
}

// Per thread range of bytes: read starts (count) and the first failing read.
// Counts are u64: a file of more than 8 GiB can hold more than 2^32 reads.
__global__ __launch_bounds__(kBlock) void k_scan_reads(const uint8_t *b, uint64_t n,
                                                       uint64_t *counts,
                                                       unsigned long long *fail) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t i0 = t * kBytesPerThread;
    uint64_t c = 0;
    if (i0 < n) {
        const uint64_t i1 = i0 + kBytesPerThread < n ? i0 + kBytesPerThread : n;
        for (uint64_t i = i0; i < i1; ++i) {
            if (!read_start(b, i)) continue;
            int32_t v;
            if (!read_value(b, n, i, v)) {
                atomicMin(fail, (unsigned long long)i);
                break;
            }
            ++c;
        }
    }
    const uint64_t nt = (n + kBytesPerThread - 1) / kBytesPerThread;
    if (t < nt) counts[t] = c;
}

// Values of the reads before the first failure, in file order.
__global__ __launch_bounds__(kBlock) void k_write_reads(const uint8_t *b, uint64_t n,
                                                        const uint64_t *offs, uint64_t fail,
                                                        int32_t *vals) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t i0 = t * kBytesPerThread;
    if (i0 >= n || i0 >= fail) return;
    const uint64_t i1 = i0 + kBytesPerThread < n ? i0 + kBytesPerThread : n;
    uint64_t k = offs[t];
    for (uint64_t i = i0; i < i1 && i < fail; ++i) {
        if (!read_start(b, i)) continue;
        int32_t v = 0;
        read_value(b, n, i, v);
        vals[k++] = v;
    }
}

// First record with a negative weight (the weighted reader stops there).
__global__ void k_first_negative(const int32_t *vals, uint64_t nrec, unsigned long long *first) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nrec;
         r += (uint64_t)gridDim.x * blockDim.x)
        if (vals[3 * r + 2] < 0) atomicMin(first, (unsigned long long)r);
}

__global__ void k_min_max(const int32_t *vals, uint64_t nrec, int arity, int *lo, int *hi) {
    int l = 2147483647, h = -2147483647 - 1;
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nrec;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const int a = vals[arity * r], c = vals[arity * r + 1];
        l = min(l, min(a, c));
        h = max(h, max(a, c));
    }
    for (int off = 32; off > 0; off >>= 1) {
        l = min(l, __shfl_xor(l, off));
        h = max(h, __shfl_xor(h, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(lo, l);
        atomicMax(hi, h);
    }
}

// Two keys per record, (x << bits | y) and (y << bits | x), value = record.
__global__ void k_edge_keys(const int32_t *vals, uint64_t nrec, int arity, int lo, int bits,
                            uint64_t *keys, uint32_t *rec) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nrec;
         r += (uint64_t)gridDim.x * blockDim.x) {
        const int64_t a = vals[arity * r], c = vals[arity * r + 1];
        const uint64_t x = (uint64_t)((a < c ? a : c) - lo), y = (uint64_t)((a < c ? c : a) - lo);
        keys[2 * r] = (x << bits) | y;
        keys[2 * r + 1] = (y << bits) | x;
        rec[2 * r] = (uint32_t)r;
        rec[2 * r + 1] = (uint32_t)r;
    }
}

// Unique (src, dst) keys: entries per key (a self-loop is entered twice).
__global__ void k_unique_entries(const uint64_t *keys, const uint32_t *rec, const uint32_t *idx,
                                 uint64_t n, int bits, uint64_t *ukey, uint32_t *urec,
                                 uint64_t *cnt) {
    const uint64_t mask = (1ull << bits) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i > 0 && keys[i] == keys[i - 1]) continue;
        const uint32_t u = idx[i] - 1;
        ukey[u] = keys[i];
        urec[u] = rec[i];
        cnt[u] = ((keys[i] >> bits) == (keys[i] & mask)) ? 2u : 1u;
    }
}

struct EmitStats {
    unsigned long long srcs, loops, pairs, wsum, non_unit;
};

// CSR from the unique keys: row starts, destinations, weights.
__global__ void k_emit_csr(const uint64_t *ukey, const uint32_t *urec, const uint64_t *pos,
                           uint64_t nu, int bits, const int32_t *vals, int arity, uint64_t *row,
                           uint32_t *col, uint32_t *w, EmitStats *st) {
    const uint64_t mask = (1ull << bits) - 1;
    unsigned long long srcs = 0, loops = 0, pairs = 0, wsum = 0, non_unit = 0;
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < nu;
         u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)(ukey[u] >> bits), y = (uint32_t)(ukey[u] & mask);
        const uint64_t p = pos[u];
        if (u == 0 || (uint32_t)(ukey[u - 1] >> bits) != x) {
            row[x] = p;
            ++srcs;
        }
        const uint32_t wt = arity == 3 ? (uint32_t)vals[3 * (uint64_t)urec[u] + 2] : 1u;
        col[p] = y;
        if (w) w[p] = wt;
        if (x == y) {
            col[p + 1] = y;
            if (w) w[p + 1] = wt;
            ++loops;
            wsum += 2ull * wt;
        } else {
            ++pairs;
            wsum += wt;
        }
        non_unit += wt != 1u;
    }
    for (int off = 32; off > 0; off >>= 1) {
        srcs += (unsigned long long)shfl_xor_u64(srcs, off);
        loops += (unsigned long long)shfl_xor_u64(loops, off);
        pairs += (unsigned long long)shfl_xor_u64(pairs, off);
        wsum += (unsigned long long)shfl_xor_u64(wsum, off);
        non_unit += (unsigned long long)shfl_xor_u64(non_unit, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&st->srcs, srcs);
        atomicAdd(&st->loops, loops);
        atomicAdd(&st->pairs, pairs);
        atomicAdd(&st->wsum, wsum);
        atomicAdd(&st->non_unit, non_unit);
    }
}

__global__ void k_degrees(const uint64_t *row, const uint32_t *w, uint32_t n, uint32_t *k,
                          unsigned long long *over) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s = row[v + 1] - row[v];
        if (w) {
            s = 0;
            for (uint64_t j = row[v]; j < row[v + 1]; ++j) s += w[j];
        }
        if (s >= 0x80000000ull) atomicAdd(over, 1ull);
        k[v] = (uint32_t)s;
    }
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (1ull << b) < x) ++b;
    return b;
}

template <typename T>
T read_back(const T *p, hipStream_t s) {
    T v;
    LV_HIP(hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

GraphDevice ingest_edge_file(const std::string &path, bool weighted, int skip_line,
                             hipStream_t s, int64_t *min_id, IngestTimes *times) {
    auto tnow = [] { return std::chrono::steady_clock::now(); };
    const auto t0 = tnow();
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(LV_ERR_IO, "cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long long fsz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    // graph.cu:22-25: skip whole lines
    uint64_t skip_off = 0;
    {
        int lines = 0;
        int ch;
        while (lines < skip_line && (ch = std::fgetc(f)) != EOF) {
            ++skip_off;
            if (ch == '\n') ++lines;
        }
        std::fseek(f, (long)skip_off, SEEK_SET);
    }
    const uint64_t n_bytes = fsz > (long long)skip_off ? (uint64_t)fsz - skip_off : 0;
    DeviceArray<uint8_t> bytes(std::max<uint64_t>(n_bytes, 1));
    {   // double-buffered pinned chunks: the read of one overlaps the copy of the other
        const size_t chunk = 64ull << 20;
        uint8_t *pin[2] = {nullptr, nullptr};
        hipEvent_t done[2];
        LV_HIP(hipHostMalloc((void **)&pin[0], chunk, hipHostMallocDefault));
        LV_HIP(hipHostMalloc((void **)&pin[1], chunk, hipHostMallocDefault));
        LV_HIP(hipEventCreate(&done[0]));
        LV_HIP(hipEventCreate(&done[1]));
        uint64_t off = 0;
        int k = 0;
        bool ok = true;
        while (off < n_bytes) {
            LV_HIP(hipEventSynchronize(done[k]));   // the buffer's previous copy finished
            const size_t len = (size_t)std::min<uint64_t>(chunk, n_bytes - off);
            if (std::fread(pin[k], 1, len, f) != len) { ok = false; break; }
            LV_HIP(hipMemcpyAsync(bytes.data() + off, pin[k], len, hipMemcpyHostToDevice, s));
            LV_HIP(hipEventRecord(done[k], s));
            off += len;
            k ^= 1;
        }
        LV_HIP(hipStreamSynchronize(s));
        (void)hipEventDestroy(done[0]);
        (void)hipEventDestroy(done[1]);
        (void)hipHostFree(pin[0]);
        (void)hipHostFree(pin[1]);
        std::fclose(f);
        if (!ok) throw Error(LV_ERR_IO, "short read of " + path);
    }
    const auto t1 = tnow();
    // reads
    const uint64_t nt = (n_bytes + kBytesPerThread - 1) / kBytesPerThread;
    DeviceArray<uint64_t> counts(nt + 1);
    DeviceArray<uint64_t> scan_tmp(scan_scratch_elems(nt + 1));
    DeviceArray<unsigned long long> fail(1);
    fill_u64(fail.data(), ~0ull, 1, s);
    if (nt) {
        const unsigned g = (unsigned)((nt + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_scan_reads, dim3(g), dim3(kBlock), 0, s, bytes.data(), n_bytes,
                           counts.data(), fail.data());
    }
    set_u64(counts.data() + nt, 0ull, s);
    scan_u64(counts.data(), counts.data(), nt + 1, false, scan_tmp.data(), s);
    const unsigned long long fail_pos = read_back(fail.data(), s);
    // reads before the first failure
    uint64_t n_reads = 0;
    {
        // counts were exclusive-scanned: the reads of threads wholly before the
        // failing byte, plus those of the failing thread before it
        const uint64_t t_fail = fail_pos == ~0ull ? nt : fail_pos / kBytesPerThread;
        n_reads = read_back(counts.data() + t_fail, s);
        if (fail_pos != ~0ull) {
            // reads of the failing thread that start before the failure
            std::vector<uint8_t> tail(kBytesPerThread + 1);
            const uint64_t i0 = t_fail * kBytesPerThread;
            const uint64_t len = std::min<uint64_t>(kBytesPerThread, n_bytes - i0);
            LV_HIP(hipMemcpyAsync(tail.data() + 1, bytes.data() + i0, len, hipMemcpyDeviceToHost, s));
            if (i0 > 0) LV_HIP(hipMemcpyAsync(tail.data(), bytes.data() + i0 - 1, 1, hipMemcpyDeviceToHost, s));
            LV_HIP(hipStreamSynchronize(s));
            auto ws = [](uint8_t c) {
                return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
            };
            auto dg = [](uint8_t c) { return c >= '0' && c <= '9'; };
            for (uint64_t i = i0; i < fail_pos; ++i) {
                const uint8_t c = tail[i - i0 + 1];
                if (ws(c)) continue;
                const bool start = i == 0 || ws(tail[i - i0]) || (dg(tail[i - i0]) && !dg(c));
                n_reads += start;
            }
        }
    }
    const int arity = weighted ? 3 : 2;
    uint64_t nrec = n_reads / arity;
    DeviceArray<int32_t> vals(std::max<uint64_t>(n_reads, 1));
    if (nt) {
        const unsigned g = (unsigned)((nt + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_write_reads, dim3(g), dim3(kBlock), 0, s, bytes.data(), n_bytes,
                           counts.data(), (uint64_t)fail_pos, vals.data());
    }
    bytes.release();
    if (weighted && nrec) {   // graph.cu: a negative weight ends the input
        fill_u64(fail.data(), ~0ull, 1, s);
        hipLaunchKernelGGL(k_first_negative, dim3(grid_for(nrec, kBlock * 4)), dim3(kBlock), 0, s,
                           vals.data(), nrec, fail.data());
        const unsigned long long neg = read_back(fail.data(), s);
        if (neg != ~0ull) nrec = neg;
    }
    if (nrec == 0) throw Error(LV_ERR_PARSE, "no edge parsed");          // graph.cu:52-55
    if (nrec > 0x7FFFFFFFull) throw Error(LV_ERR_RANGE, "more than 2^31-1 edges");
    const auto t2 = tnow();
    DeviceArray<int> lohi(2);
    {
        const int init[2] = {2147483647, -2147483647 - 1};
        LV_HIP(hipMemcpyAsync(lohi.data(), init, sizeof(init), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_min_max, dim3(grid_for(nrec, kBlock * 4)), dim3(kBlock), 0, s,
                           vals.data(), nrec, arity, lohi.data(), lohi.data() + 1);
    }
    int lh[2];
    LV_HIP(hipMemcpyAsync(lh, lohi.data(), sizeof(lh), hipMemcpyDeviceToHost, s));
    LV_HIP(hipStreamSynchronize(s));
    const int64_t lo = lh[0], hi = lh[1];
    if (hi - lo >= (int64_t)0xFFFFFFFFll) throw Error(LV_ERR_RANGE, "id span >= 2^32-1");
    const uint64_t span = (uint64_t)(hi - lo) + 1;
    const int bits = std::max(bits_for(span), 1);
    const uint64_t nk = 2 * nrec;
    DeviceArray<uint64_t> k0(nk), k1(nk);
    DeviceArray<uint32_t> v0(nk), v1(nk);
    hipLaunchKernelGGL(k_edge_keys, dim3(grid_for(nrec, kBlock * 4)), dim3(kBlock), 0, s,
                       vals.data(), nrec, arity, (int)lo, bits, k0.data(), v0.data());
    {
        DeviceArray<uint32_t> hist(radix_hist_elems(nk));
        DeviceArray<uint32_t> hscan(scan_scratch_elems(radix_hist_elems(nk)));
        radix_sort_pairs(k0.data(), v0.data(), k1.data(), v1.data(), nk, 2 * bits, hist.data(),
                         hscan.data(), s);   // stable: records in file order within a key
    }
    DeviceArray<uint32_t> idx(nk);
    {
        DeviceArray<uint32_t> tmp(scan_scratch_elems(nk));
        heads_u64(k0.data(), nk, idx.data(), s);
        scan_u32(idx.data(), idx.data(), nk, true, tmp.data(), s);
    }
    const uint32_t nu = read_back(idx.data() + (nk - 1), s);
    DeviceArray<uint64_t> ukey(nu), pos((uint64_t)nu + 1);
    DeviceArray<uint32_t> urec(nu);
    hipLaunchKernelGGL(k_unique_entries, dim3(grid_for(nk, kBlock * 4)), dim3(kBlock), 0, s,
                       k0.data(), v0.data(), idx.data(), nk, bits, ukey.data(), urec.data(),
                       pos.data());
    set_u64(pos.data() + nu, 0ull, s);
    {
        DeviceArray<uint64_t> tmp(scan_scratch_elems((uint64_t)nu + 1));
        scan_u64(pos.data(), pos.data(), (uint64_t)nu + 1, false, tmp.data(), s);
    }
    const uint64_t E = read_back(pos.data() + nu, s);
    k0.release(); k1.release(); v0.release(); v1.release(); idx.release();
    GraphDevice g;
    g.n_nodes = (uint32_t)span;
    g.n_entries = E;
    g.neighboorhood_sum.resize(span + 1);
    g.edge_destination.resize(std::max<uint64_t>(E, 1));
    DeviceArray<uint32_t> wbuf(weighted ? std::max<uint64_t>(E, 1) : 1);
    DeviceArray<EmitStats> est(1);
    fill_u64(reinterpret_cast<unsigned long long *>(est.data()), 0ull,
             sizeof(EmitStats) / sizeof(unsigned long long), s);
    hipLaunchKernelGGL(k_emit_csr, dim3(grid_for(nu, kBlock * 4)), dim3(kBlock), 0, s,
                       ukey.data(), urec.data(), pos.data(), (uint64_t)nu, bits, vals.data(), arity,
                       g.neighboorhood_sum.data(), g.edge_destination.data(),
                       weighted ? wbuf.data() : nullptr, est.data());
    set_u64(g.neighboorhood_sum.data() + span, E, s);
    const EmitStats es = read_back(est.data(), s);
    if (es.srcs != span) throw Error(LV_ERR_NOT_DENSE, "vertex ids are not dense");
    g.unit_weights = es.non_unit == 0;
    if (!g.unit_weights) g.weights = std::move(wbuf);
    g.n_links = es.loops + es.pairs / 2;
    g.total_weight = (double)es.wsum / 2;                               // graph.cu:81
    g.m2 = es.wsum;
    if (g.m2 >= 0x80000000ull) throw Error(LV_ERR_RANGE, "total weight 2m >= 2^31");
    g.tot_weight_per_nodes.resize(std::max<uint64_t>(span, 1));
    DeviceArray<unsigned long long> over(1);
    fill_u64(over.data(), 0ull, 1, s);
    hipLaunchKernelGGL(k_degrees, dim3(grid_for(span, kBlock * 4)), dim3(kBlock), 0, s,
                       g.neighboorhood_sum.data(), g.unit_weights ? nullptr : g.weights.data(),
                       (uint32_t)span, g.tot_weight_per_nodes.data(), over.data());
    if (read_back(over.data(), s)) throw Error(LV_ERR_RANGE, "vertex degree >= 2^31");
    if (min_id) *min_id = lo;
    if (times) {
        const auto t3 = tnow();
        times->read_upload = std::chrono::duration<double>(t1 - t0).count();
        times->tokenize = std::chrono::duration<double>(t2 - t1).count();
        times->build = std::chrono::duration<double>(t3 - t2).count();
    }
    return g;
}

}  // namespace lv
