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

int scratch_alloc(void** p, size_t bytes, hipStream_t s) {
    hipMemPool_t pool = scratch_pool();
    const hipError_t e = pool ? hipMallocFromPoolAsync(p, bytes, pool, s) : hipMallocAsync(p, bytes, s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(PS_ERR_OUT_OF_MEMORY, "device scratch of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
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
