/*
 * device_array.h -- owning HIP device buffer and the library's error type.
 * Replaces the reference's thrust::device_vector members (community.h:16-20,
 * graph.cuh:25-32) with plain hipMalloc'd storage; allocation failure throws
 * lv::Error(LV_ERR_OOM) the way thrust throws std::bad_alloc.
 */
#ifndef LV_DEVICE_ARRAY_H
#define LV_DEVICE_ARRAY_H

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "lv_api.h"

namespace lv {

struct Error : std::runtime_error {
    lv_status status;
    Error(lv_status s, const std::string &what) : std::runtime_error(what), status(s) {}
};

inline void hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return;
    const lv_status s = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? LV_ERR_OOM
                        : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? LV_ERR_NO_DEVICE
                                                                                 : LV_ERR_HIP;
    throw Error(s, std::string(what) + ": " + hipGetErrorString(e));
}

#define LV_HIP(call) ::lv::hip_check((call), #call)

// Waits for the work queued on stream s.  A run's host thread has nothing to
// do while it waits for list sizes or level sizes (~30 times per run, for
// tens of microseconds each), so it polls the stream rather than sleeping in
// a blocking wait; after 20 ms of polling it blocks.
inline void wait_stream(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) hip_check(e, "hipStreamQuery");
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
            return;
        }
    }
}

template <typename T>
class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_), owned_(o.owned_) {
        o.p_ = nullptr; o.n_ = 0; o.owned_ = true;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; owned_ = o.owned_;
            o.p_ = nullptr; o.n_ = 0; o.owned_ = true;
        }
        return *this;
    }
    // Non-owning view of another array's storage (never freed by this object).
    static DeviceArray borrow(const DeviceArray &o) {
        DeviceArray v;
        v.p_ = const_cast<T *>(o.p_);
        v.n_ = o.n_;
        v.owned_ = false;
        return v;
    }
    ~DeviceArray() { release(); }

    void resize(size_t n) {
        release();
        if (n == 0) return;
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, n * sizeof(T));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(e == hipErrorOutOfMemory ? LV_ERR_OOM : LV_ERR_HIP,
                        std::string("hipMalloc(") + std::to_string(n * sizeof(T)) + " bytes): " +
                            hipGetErrorString(e));
        }
        p_ = static_cast<T *>(q);
        n_ = n;
        owned_ = true;
    }
    // Keep the storage when it already holds n elements (grow-only reuse).
    void ensure(size_t n) {
        if (n <= n_ && owned_) return;
        resize(n);
    }
    void release() {
        if (p_ && owned_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
        owned_ = true;
    }
    T *data() { return p_; }
    const T *data() const { return p_; }
    size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
    bool owned_ = true;
};

}  // namespace lv

#endif
