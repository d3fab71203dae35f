// common.hpp — error plumbing for the C-ABI: HIP failures become std::runtime_error inside
// the library and a nonzero status + thread-local message at the boundary (the reference's
// engines throw "CUDA <action>: <msg>", src/common/CudaSupport.cuh:6-9).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

namespace swl {
std::string &last_error();
// Held while an engine is built (allocations, synchronous uploads) and while a stream is captured into a
// hipGraph, so that synchronous HIP calls of a session being built on another thread cannot invalidate
// a capture in progress.
std::mutex &device_setup_mutex();

inline void check(hipError_t e, const char *action) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP ") + action + ": " + hipGetErrorString(e));
}
inline void require(bool ok, const char *message) {
    if (!ok) throw std::runtime_error(message);
}

// Runs f(); converts exceptions to an int status and records the message.
template <class F> int guarded(F &&f) {
    try {
        f();
        last_error().clear();
        return 0;
    } catch (const std::exception &e) {
        last_error() = e.what();
        return 1;
    } catch (...) {
        last_error() = "unknown error";
        return 1;
    }
}

// RAII device allocation (zero-filled).
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr, o.bytes = 0; }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        std::swap(ptr, o.ptr);
        std::swap(bytes, o.bytes);
        return *this;
    }
    ~DeviceBuffer() {
        if (ptr) (void)hipFree(ptr);
    }
    void allocate(size_t n) {
        if (ptr) (void)hipFree(ptr), ptr = nullptr;
        bytes = n;
        if (!n) return;
        check(hipMalloc(&ptr, n), "allocation");
        // the clear runs on the null stream, which the engines' non-blocking streams do not wait for:
        // finish it here so a kernel enqueued next on such a stream cannot run before it
        check(hipMemsetAsync(ptr, 0, n, nullptr), "clear allocation");
        check(hipStreamSynchronize(nullptr), "clear allocation");
    }
    void upload(const void *host, size_t n) {
        allocate(n);
        if (n) check(hipMemcpy(ptr, host, n, hipMemcpyHostToDevice), "upload");
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Switches the current device for the scope (DeviceScope, CudaSupport.cuh:10-24).
struct DeviceScope {
    int previous = -1;
    explicit DeviceScope(int device) {
        check(hipGetDevice(&previous), "query device");
        if (previous != device) check(hipSetDevice(device), "select device");
    }
    ~DeviceScope() {
        if (previous >= 0) (void)hipSetDevice(previous);
    }
};
} // namespace swl
