// common.hpp — error plumbing for the C-ABI: HIP failures become std::runtime_error inside
// the library and a nonzero status + thread-local message at the boundary (the reference's
// engines throw "CUDA <action>: <msg>", src/common/CudaSupport.cuh:6-9).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#include <algorithm>

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
// Non-blocking stream of the calling thread for set-up work on the current device (default priority).  Clears,
// uploads and set-up kernels go here and the thread waits for this stream only: a session built on one thread
// while another keeps the device busy does not wait for that work (as a null-stream synchronisation would).
// The streams live until the process ends.
inline hipStream_t setup_stream() {
    thread_local std::vector<hipStream_t> streams;
    int device = 0;
    check(hipGetDevice(&device), "query device");
    if ((int)streams.size() <= device) streams.resize(device + 1, nullptr);
    if (!streams[device]) check(hipStreamCreateWithFlags(&streams[device], hipStreamNonBlocking), "set-up stream");
    return streams[device];
}

// Page-locked staging of the calling thread for set-up uploads (copies from it are DMA transfers that need no
// runtime staging).  Blocks never move, so copies may still be reading earlier regions while later ones are
// taken; setup_finish() waits for the set-up stream and recycles the regions.  Blocks are kept for the life of
// the process (page-locking is paid once per thread).
struct PinnedArena {
    struct Block {
        char *p;
        size_t size, used;
    };
    std::vector<Block> blocks;
    void *take(size_t n) {
        n = (n + 255) & ~size_t(255);
        for (auto &b : blocks)
            if (b.size - b.used >= n) {
                void *r = b.p + b.used;
                b.used += n;
                return r;
            }
        const size_t size = std::max<size_t>(n, size_t(64) << 20);
        char *p = nullptr;
        check(hipHostMalloc(reinterpret_cast<void **>(&p), size, hipHostMallocDefault), "page-locked staging");
        blocks.push_back({p, size, n});
        return p;
    }
    void reset() {
        for (auto &b : blocks) b.used = 0;
    }
};
inline PinnedArena &pinned_arena() {
    thread_local PinnedArena arena;
    return arena;
}
// Completes every set-up copy and clear the calling thread has enqueued (upload_async / allocate_async).
inline void setup_finish() {
    check(hipStreamSynchronize(setup_stream()), "set-up");
    pinned_arena().reset();
}

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
        // cleared on this thread's set-up stream and finished here, so that a kernel enqueued next on any
        // other stream (the engines' are non-blocking) sees zeros
        const hipStream_t s = setup_stream();
        check(hipMemsetAsync(ptr, 0, n, s), "clear allocation");
        check(hipStreamSynchronize(s), "clear allocation");
    }
    void upload(const void *host, size_t n) {
        if (ptr) (void)hipFree(ptr), ptr = nullptr;
        bytes = n;
        if (!n) return;
        check(hipMalloc(&ptr, n), "allocation");
        const hipStream_t s = setup_stream();
        check(hipMemcpyAsync(ptr, host, n, hipMemcpyHostToDevice, s), "upload");
        check(hipStreamSynchronize(s), "upload");
    }
    // Set-up variants: enqueued on the set-up stream without waiting; setup_finish() completes them before the
    // memory is used from any other stream.
    void allocate_async(size_t n) {
        if (ptr) (void)hipFree(ptr), ptr = nullptr;
        bytes = n;
        if (!n) return;
        check(hipMalloc(&ptr, n), "allocation");
        check(hipMemsetAsync(ptr, 0, n, setup_stream()), "clear allocation");
    }
    void upload_staged(const void *pinned, size_t n) { // `pinned` from pinned_arena()
        if (ptr) (void)hipFree(ptr), ptr = nullptr;
        bytes = n;
        if (!n) return;
        check(hipMalloc(&ptr, n), "allocation");
        check(hipMemcpyAsync(ptr, pinned, n, hipMemcpyHostToDevice, setup_stream()), "upload");
    }
    void upload_async(const void *host, size_t n) {
        void *s = n ? pinned_arena().take(n) : nullptr;
        if (n && host) std::memcpy(s, host, n);
        else if (n) std::memset(s, 0, n);
        upload_staged(s, n);
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Temporary device memory of a set-up routine, stream-ordered on the set-up stream: released without the
// device-wide synchronisation of hipFree, so building a session never waits for other streams' work.
struct SetupBuffer {
    void *ptr = nullptr;
    hipStream_t stream = nullptr;
    SetupBuffer() = default;
    SetupBuffer(const SetupBuffer &) = delete;
    SetupBuffer &operator=(const SetupBuffer &) = delete;
    ~SetupBuffer() {
        if (ptr) (void)hipFreeAsync(ptr, stream);
    }
    void allocate(size_t n) {
        stream = setup_stream();
        check(hipMallocAsync(&ptr, std::max<size_t>(n, 16), stream), "set-up allocation");
        check(hipMemsetAsync(ptr, 0, std::max<size_t>(n, 16), stream), "set-up clear"); // ordered before its use
    }
    void upload(const void *host, size_t n) {
        allocate(n);
        if (n) check(hipMemcpyAsync(ptr, host, n, hipMemcpyHostToDevice, stream), "set-up upload");
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
