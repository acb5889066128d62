// Runtime part of the C ABI: device selection, streams, memory, copies and
// error reporting.  Replaces the reference's CUDA shim (Puzzles/GPU.cu:5-58),
// whose ERR() throws; here every failure becomes a status code plus a
// per-thread message.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <vector>

#include "ps_internal.h"

namespace ps {

namespace {
thread_local std::string t_last_error;
std::mutex g_init_mutex;
bool g_ready = false;
int g_device = -1;
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
    return code;
}

void clear_error() { t_last_error.clear(); }

bool runtime_ready() { return g_ready; }

int require_runtime() {
    if (!g_ready) return fail(PS_ERR_NOT_INITIALIZED, "ps_init() has not been called");
    return PS_OK;
}

__global__ void k_fill_bytes(uint8_t* dst, uint8_t v, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // bulk: 16-byte stores on the aligned part
    const uint64_t head = (16 - ((uintptr_t)dst & 15)) & 15;
    const uint64_t h = head < n ? head : n;
    if (i0 < h) dst[i0] = v;
    const uint64_t nvec = (n - h) / 16;
    const uint32_t w = 0x01010101u * v;
    uint4* vd = reinterpret_cast<uint4*>(dst + h);
    for (uint64_t i = i0; i < nvec; i += stride) vd[i] = make_uint4(w, w, w, w);
    const uint64_t tail0 = h + nvec * 16;
    if (i0 < n - tail0) dst[tail0 + i0] = v;
}

int hip_cu_count() {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) {
        (void)hipGetLastError();
        return 256;
    }
    return v;
}

// The library's stream-ordered scratch pool on the current device.  Freed
// blocks stay reserved in it up to 1 GiB, so the short-lived buffers of the
// codec and sorting helpers are taken from the pool instead of being mapped
// again on every call.
hipMemPool_t scratch_pool() {
    constexpr int MAX_DEV = 64;
    static hipMemPool_t pools[MAX_DEV] = {};
    static std::mutex m;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(m);
    if (!pools[dev]) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        uint64_t keep = 1ull << 30;
        if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) (void)hipGetLastError();
        pools[dev] = pool;
    }
    return pools[dev];
}

namespace {
// Small scratch blocks released by this thread, kept for reuse by this thread
// on the stream they were released on: work queued later on that stream runs
// after the work that used the block, as with hipFreeAsync followed by
// hipMallocAsync on one stream.  hipFreeAsync costs about 10 us of host time
// per call, which the short-lived helper buffers of a segment step would pay
// a few dozen times per step.
struct CachedBlock {
    hipStream_t s;
    int dev;
    void* p;
    size_t bytes;
};
constexpr size_t CACHE_MAX_BLOCK = 1u << 20;
constexpr size_t CACHE_MAX_BLOCKS = 64;
thread_local std::vector<CachedBlock> t_cache;
}  // namespace

void scratch_free(void* p, size_t bytes, hipStream_t s) {
    if (!p) return;
    int dev = 0;
    if (bytes <= CACHE_MAX_BLOCK && t_cache.size() < CACHE_MAX_BLOCKS && hipGetDevice(&dev) == hipSuccess) {
        t_cache.push_back(CachedBlock{s, dev, p, bytes});
        return;
    }
    (void)hipFreeAsync(p, s);
}

int scratch_alloc(void** p, size_t bytes, hipStream_t s) {
    if (bytes <= CACHE_MAX_BLOCK && !t_cache.empty()) {
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            size_t best = t_cache.size();
            for (size_t i = 0; i < t_cache.size(); ++i) {
                const CachedBlock& b = t_cache[i];
                if (b.s == s && b.dev == dev && b.bytes >= bytes && b.bytes <= std::max<size_t>(2 * bytes, 4096) &&
                    (best == t_cache.size() || b.bytes < t_cache[best].bytes))
                    best = i;
            }
            if (best < t_cache.size()) {
                *p = t_cache[best].p;
                t_cache[best] = t_cache.back();
                t_cache.pop_back();
                return PS_OK;
            }
        }
    }
    hipMemPool_t pool = scratch_pool();
    const hipError_t e = pool ? hipMallocFromPoolAsync(p, bytes, pool, s) : hipMallocAsync(p, bytes, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(PS_ERR_OUT_OF_MEMORY, "device scratch of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    return PS_OK;
}

namespace {
// Pinned staging for small host-to-device copies: a pageable source costs a
// synchronous staging copy inside hipMemcpyAsync (about 60 us for 128 KB);
// from pinned memory the copy is queued and the call returns at once.  A
// ring of four buffers per thread, each reused only after the event recorded
// behind its last copy has passed.
struct Stage {
    void* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    int dev = -1;
    bool pending = false;
};
thread_local Stage t_stage[4];
thread_local int t_stage_next = 0;
}  // namespace

int copy_to_device(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return PS_OK;
    if (bytes > STAGE_MAX_BYTES) {
        PS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        return PS_OK;
    }
    int dev = 0;
    PS_HIP_CHECK(s ? hipStreamGetDevice(s, &dev) : hipGetDevice(&dev));
    Stage& st = t_stage[t_stage_next];
    t_stage_next = (t_stage_next + 1) % 4;
    if (st.pending) {
        hipError_t q;
        while ((q = hipEventQuery(st.ev)) == hipErrorNotReady) {
        }
        if (q != hipSuccess) return fail(PS_ERR_HIP, "waiting for a staged copy failed: %s", hipGetErrorString(q));
        st.pending = false;
    }
    if (st.ev && st.dev != dev) {
        (void)hipEventDestroy(st.ev);
        st.ev = nullptr;
    }
    if (!st.ev) {
        int cur = 0;
        PS_HIP_CHECK(hipGetDevice(&cur));
        if (cur != dev) PS_HIP_CHECK(hipSetDevice(dev));
        const hipError_t e = hipEventCreateWithFlags(&st.ev, hipEventDisableTiming);
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) {
            st.ev = nullptr;
            return fail(PS_ERR_HIP, "event creation failed: %s", hipGetErrorString(e));
        }
        st.dev = dev;
    }
    if (st.cap < bytes) {
        if (st.p) (void)hipHostFree(st.p);
        st.p = nullptr;
        st.cap = 0;
        const size_t want = std::max<size_t>(bytes, 64 * 1024);
        if (hipHostMalloc(&st.p, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            st.p = nullptr;
            return fail(PS_ERR_OUT_OF_MEMORY, "pinned staging of %zu bytes", want);
        }
        st.cap = want;
    }
    std::memcpy(st.p, src, bytes);
    PS_HIP_CHECK(hipMemcpyAsync(dst, st.p, bytes, hipMemcpyHostToDevice, s));
    PS_HIP_CHECK(hipEventRecord(st.ev, s));
    st.pending = true;
    return PS_OK;
}

size_t hip_lds_limit() {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
        (void)hipGetLastError();
        return 65536;
    }
    return (size_t)v;
}

// 16-byte vector copy; both pointers and the size must be multiples of 16
__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

int launch_fill(void* dst, uint8_t value, uint64_t bytes, hipStream_t s) {
    if (bytes == 0) return PS_OK;
    uint64_t blocks = (bytes / 16 + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_bytes, dim3((uint32_t)blocks), dim3(256), 0, s, (uint8_t*)dst, value, bytes);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

}  // namespace ps

using namespace ps;

extern "C" {

const char* ps_last_error(void) { return t_last_error.c_str(); }

int ps_device_count(int32_t* count) {
    PS_API_BEGIN
    if (!count) return fail(PS_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return PS_OK;
    PS_API_END
}

int ps_init(int32_t device) {
    PS_API_BEGIN
    std::lock_guard<std::mutex> lock(g_init_mutex);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(PS_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(PS_ERR_INVALID_ARGUMENT, "device %d out of range [0,%d)", device, n);
    PS_HIP_CHECK(hipSetDevice(device));
    PS_HIP_CHECK(hipFree(nullptr));  // force context creation
    if (int r = bits_scratch_init(device)) return r;
    g_device = device;
    g_ready = true;
    clear_error();
    return PS_OK;
    PS_API_END
}

int ps_shutdown(void) {
    PS_API_BEGIN
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_ready) PS_HIP_CHECK(hipDeviceSynchronize());
    g_ready = false;
    return PS_OK;
    PS_API_END
}

int ps_device_props(char* name, int32_t name_cap, int32_t* cu_count, int32_t* lds_bytes_per_block,
                    uint64_t* hbm_bytes) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    hipDeviceProp_t prop;
    PS_HIP_CHECK(hipGetDeviceProperties(&prop, g_device));
    if (name && name_cap > 0) {
        snprintf(name, (size_t)name_cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (lds_bytes_per_block) *lds_bytes_per_block = (int32_t)prop.sharedMemPerBlock;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return PS_OK;
    PS_API_END
}

int ps_stream_create(ps_stream* out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!out) return fail(PS_ERR_INVALID_ARGUMENT, "out is NULL");
    hipStream_t s;
    PS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = reinterpret_cast<ps_stream>(s);
    return PS_OK;
    PS_API_END
}

int ps_stream_destroy(ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!s) return PS_OK;
    PS_HIP_CHECK(hipStreamDestroy(stream_of(s)));
    return PS_OK;
    PS_API_END
}

int ps_stream_sync(ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    PS_HIP_CHECK(hipStreamSynchronize(stream_of(s)));
    return PS_OK;
    PS_API_END
}

int ps_device_sync(void) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    PS_HIP_CHECK(hipDeviceSynchronize());
    return PS_OK;
    PS_API_END
}

int ps_alloc(void** out, uint64_t bytes) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!out) return fail(PS_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (bytes == 0) bytes = 16;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        return fail(PS_ERR_OUT_OF_MEMORY, "hipMalloc(%llu bytes): out of device memory", (unsigned long long)bytes);
    }
    PS_HIP_CHECK(e);
    *out = p;
    return PS_OK;
    PS_API_END
}

int ps_free(void* p) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p) return PS_OK;
    PS_HIP_CHECK(hipFree(p));
    return PS_OK;
    PS_API_END
}

int ps_host_alloc(void** out, uint64_t bytes) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!out) return fail(PS_ERR_INVALID_ARGUMENT, "out is NULL");
    void* p = nullptr;
    PS_HIP_CHECK(hipHostMalloc(&p, bytes ? bytes : 16, hipHostMallocDefault));
    *out = p;
    return PS_OK;
    PS_API_END
}

int ps_host_free(void* p) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p) return PS_OK;
    PS_HIP_CHECK(hipHostFree(p));
    return PS_OK;
    PS_API_END
}

int ps_memcpy_h2d_async(void* dst, const void* src, uint64_t bytes, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bytes == 0) return PS_OK;
    if (!dst || !src) return fail(PS_ERR_INVALID_ARGUMENT, "NULL pointer");
    PS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(s)));
    return PS_OK;
    PS_API_END
}

int ps_memcpy_d2h_async(void* dst, const void* src, uint64_t bytes, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bytes == 0) return PS_OK;
    if (!dst || !src) return fail(PS_ERR_INVALID_ARGUMENT, "NULL pointer");
    PS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(s)));
    return PS_OK;
    PS_API_END
}

int ps_memcpy_d2d_async(void* dst, const void* src, uint64_t bytes, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bytes == 0) return PS_OK;
    if (!dst || !src) return fail(PS_ERR_INVALID_ARGUMENT, "NULL pointer");
    PS_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(s)));
    return PS_OK;
    PS_API_END
}

int ps_fill_async(void* dst, uint8_t value, uint64_t bytes, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bytes && !dst) return fail(PS_ERR_INVALID_ARGUMENT, "NULL pointer");
    return launch_fill(dst, value, bytes, stream_of(s));
    PS_API_END
}

int ps_copy_async(void* dst, const void* src, uint64_t bytes, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bytes == 0) return PS_OK;
    if (!dst || !src) return fail(PS_ERR_INVALID_ARGUMENT, "NULL pointer");
    if (((uintptr_t)dst | (uintptr_t)src | bytes) & 15)
        return fail(PS_ERR_INVALID_ARGUMENT, "pointers and size must be multiples of 16 bytes");
    const uint64_t n16 = bytes / 16;
    const uint64_t blocks = std::min<uint64_t>((n16 + 255) / 256, (uint64_t)hip_cu_count() * 16);
    hipLaunchKernelGGL(k_copy16, dim3((uint32_t)std::max<uint64_t>(1, blocks)), dim3(256), 0, stream_of(s),
                       (uint4*)dst, (const uint4*)src, n16);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

int ps_puzzle_query(const ps_puzzle* p, ps_puzzle_info* out) {
    PS_API_BEGIN
    return puzzle_query(p, out);
    PS_API_END
}

int ps_expand_slots(const ps_puzzle* p, int32_t mode) {
    try {
        return slots_per_parent(p, mode);
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
