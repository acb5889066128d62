// Scratch device buffers and host<->device vector copies for the host-side
// orchestration code (codecs, segment step).  Allocation is stream-ordered
// (hipMallocAsync / hipFreeAsync on the owning stream).
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
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
    int alloc(size_t bytes, hipStream_t st) {
        s = st;
        hipError_t e = hipMallocAsync(&p, bytes ? bytes : 16, st);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(PS_ERR_OUT_OF_MEMORY, "device scratch of %zu bytes: %s", bytes, hipGetErrorString(e));
        }
        return PS_OK;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

template <class T>
inline int to_device(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
    if (int r = b.alloc(v.size() * sizeof(T), s)) return r;
    if (!v.empty()) PS_HIP_CHECK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return PS_OK;
}

template <class T>
inline int to_host(std::vector<T>& v, const void* d, size_t n, hipStream_t s) {
    v.resize(n);
    if (n) PS_HIP_CHECK(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
}

inline uint32_t blocks_for(uint64_t n, uint32_t tpb) {
    return (uint32_t)std::max<uint64_t>(1, (n + tpb - 1) / tpb);
}

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
        hipError_t e = hipMallocAsync(&p, want, s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(PS_ERR_OUT_OF_MEMORY, "segment step buffer of %zu bytes: %s", want, hipGetErrorString(e));
        }
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
