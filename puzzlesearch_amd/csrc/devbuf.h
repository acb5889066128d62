// Scratch device buffers and host<->device vector copies for the host-side
// orchestration code (codecs, segment step).  Allocation is stream-ordered,
// from the library's scratch pool, and freed with hipFreeAsync on the owning
// stream.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ps_internal.h"

namespace ps {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { scratch_free(p, n, s); }
    int alloc(size_t bytes, hipStream_t st) {
        scratch_free(p, n, s);
        p = nullptr;
        s = st;
        n = bytes ? bytes : 16;
        return scratch_alloc(&p, n, st);
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

// Polls `e` until the work before it has finished (see stream_wait).
inline int event_wait(hipEvent_t e) {
    hipError_t q;
    while ((q = hipEventQuery(e)) == hipErrorNotReady) {
    }
    if (q != hipSuccess) return fail(PS_ERR_HIP, "waiting for the stream failed: %s", hipGetErrorString(q));
    return PS_OK;
}

// Waits until the work queued on `s` so far has finished, polling an event
// from this thread instead of sleeping in hipStreamSynchronize: the steps that
// read a count back between launches resume within microseconds of the copy
// instead of after an interrupt wake-up (measured 30-390 us per wait).  One
// event per thread and device, kept for the life of the thread.
inline int stream_wait(hipStream_t s) {
    constexpr int MAX_DEV = 64;
    thread_local hipEvent_t ev[MAX_DEV] = {};
    int dev = 0;
    PS_HIP_CHECK(s ? hipStreamGetDevice(s, &dev) : hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEV) return fail(PS_ERR_HIP, "device %d out of range", dev);
    if (!ev[dev]) {
        int cur = 0;
        PS_HIP_CHECK(hipGetDevice(&cur));
        if (cur != dev) PS_HIP_CHECK(hipSetDevice(dev));
        const hipError_t e = hipEventCreateWithFlags(&ev[dev], hipEventDisableTiming);
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) {
            ev[dev] = nullptr;
            return fail(PS_ERR_HIP, "event creation failed: %s", hipGetErrorString(e));
        }
    }
    PS_HIP_CHECK(hipEventRecord(ev[dev], s));
    return event_wait(ev[dev]);
}

template <class T>
inline int to_device(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    if (int r = b.alloc(v.size() * sizeof(T), s)) return r;
    return copy_to_device(b.p, v.data(), v.size() * sizeof(T), s);
}

template <class T>
inline int to_host(std::vector<T>& v, const void* d, size_t n, hipStream_t s) {
    v.resize(n);
    if (n) PS_HIP_CHECK(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    if (int r = stream_wait(s)) return r;
    return PS_OK;
}

inline uint32_t blocks_for(uint64_t n, uint32_t tpb) {
    return (uint32_t)std::max<uint64_t>(1, (n + tpb - 1) / tpb);
}

// A grow-only pinned host buffer of a workspace, the target of the
// read-backs that complete with the step; regrown only between steps, when
// no copy into it is pending.
struct GrowHost {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap && p) return PS_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(PS_ERR_OUT_OF_MEMORY, "pinned host buffer of %zu bytes", want);
        }
        cap = want;
        return PS_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

// A grow-only device buffer of a workspace (segment steps): reallocated on
// its stream, with a quarter to spare, only when a step needs more.
struct GrowBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes, hipStream_t s) {
        if (bytes <= cap && p) return PS_OK;
        if (p) (void)hipFreeAsync(p, s);
        p = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
        if (int r = scratch_alloc(&p, want, s)) return r;
        cap = want;
        return PS_OK;
    }
    void release(hipStream_t s) {
        if (p) (void)hipFreeAsync(p, s);
        p = nullptr;
        cap = 0;
    }
    // a buffer handed to a kernel must have been sized by ensure() first
    template <class T>
    T* as() const {
        if (!p) {
            std::fprintf(stderr, "puzzlesearch_amd: segment step used a buffer before sizing it\n");
            std::abort();
        }
        return reinterpret_cast<T*>(p);
    }
    // the buffer for reporting to the caller: NULL while it was never sized (nothing in it)
    template <class T>
    T* ptr_or_null() const {
        return reinterpret_cast<T*>(p);
    }
};

}  // namespace ps
