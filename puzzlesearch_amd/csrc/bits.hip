// Device bit-array operations of the C ABI (Common/BitArray.h:69-169 on HBM).
#include <mutex>

#include "bfs_kernels.h"
#include "ps_internal.h"

namespace ps {

namespace {

constexpr int TPB = 256;

uint32_t grid_for(uint64_t items, uint64_t per_thread = 1) {
    uint64_t b = (items + (uint64_t)TPB * per_thread - 1) / ((uint64_t)TPB * per_thread);
    if (b < 1) b = 1;
    if (b > 16384) b = 16384;
    return (uint32_t)b;
}

template <int T>
__device__ __forceinline__ void block_atomic_add(unsigned long long v, unsigned long long* g) {
    __shared__ unsigned long long red[T / 64];
    v = psk::wave_sum(v);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < T / 64; ++w) s += red[w];
        if (s) atomicAdd(g, s);
    }
}

__global__ __launch_bounds__(TPB) void k_andnot_count(uint64_t* a, const uint64_t* e1, const uint64_t* e2,
                                                      uint64_t n, unsigned long long* count) {
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * TPB) {
        uint64_t v = a[i] & ~e1[i];
        if (e2) v &= ~e2[i];
        a[i] = v;
        c += (unsigned long long)__popcll(v);
    }
    block_atomic_add<TPB>(c, count);
}

__global__ __launch_bounds__(TPB) void k_count(const uint64_t* a, uint64_t n, unsigned long long* count) {
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * TPB)
        c += (unsigned long long)__popcll(a[i]);
    block_atomic_add<TPB>(c, count);
}

__global__ __launch_bounds__(TPB) void k_hash(const uint64_t* a, uint64_t n, uint64_t off, unsigned long long* h) {
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * TPB)
        s += psk::word_hash(i + off, a[i]);
    block_atomic_add<TPB>(s, h);
}

// Indexes >= nbits are not written; their number goes to *rejected (if given).
__global__ __launch_bounds__(TPB) void k_set_list(uint32_t* a, uint64_t nbits, const uint64_t* idx, uint64_t n,
                                                  unsigned long long* rejected) {
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * TPB) {
        const uint64_t v = idx[i];
        if (v < nbits) psk::gor(&a[v >> 5], 1u << (v & 31));
        else ++bad;
    }
    if (rejected) block_atomic_add<TPB>(bad, rejected);
}

// ScanBits into a list, in three passes over chunks of SCAN_CHUNK words:
// per-chunk popcounts, one exclusive scan, per-chunk ordered writes.
constexpr uint64_t SCAN_CHUNK = 4096;

__global__ __launch_bounds__(TPB) void k_scan_counts(const uint64_t* a, uint64_t nwords, uint64_t nbits,
                                                     unsigned long long* chunk_counts) {
    const uint64_t c = blockIdx.x;
    const uint64_t w0 = c * SCAN_CHUNK;
    uint64_t w1 = w0 + SCAN_CHUNK;
    if (w1 > nwords) w1 = nwords;
    unsigned long long s = 0;
    for (uint64_t i = w0 + threadIdx.x; i < w1; i += TPB) {
        uint64_t v = a[i];
        if (i == nwords - 1 && (nbits & 63)) v &= (1ull << (nbits & 63)) - 1ull;
        s += (unsigned long long)__popcll(v);
    }
    s = psk::wave_sum(s);
    __shared__ unsigned long long red[TPB / 64];
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < TPB / 64; ++w) t += red[w];
        chunk_counts[c] = t;
    }
}

// exclusive scan of n values in one block; total in out[n]
__global__ __launch_bounds__(1024) void k_scan_excl(unsigned long long* v, uint64_t n) {
    __shared__ unsigned long long part[1024];
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = threadIdx.x * per, e = (b + per < n) ? b + per : n;
    unsigned long long s = 0;
    for (uint64_t i = b; i < e; ++i) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long run = 0;
        for (int i = 0; i < 1024; ++i) {
            const unsigned long long t = part[i];
            part[i] = run;
            run += t;
        }
        v[n] = run;
    }
    __syncthreads();
    unsigned long long run = part[threadIdx.x];
    for (uint64_t i = b; i < e; ++i) {
        const unsigned long long t = v[i];
        v[i] = run;
        run += t;
    }
}

__global__ __launch_bounds__(TPB) void k_scan_write(const uint64_t* a, uint64_t nwords, uint64_t nbits,
                                                    const unsigned long long* chunk_off, uint64_t* out, uint64_t cap) {
    const uint64_t c = blockIdx.x;
    const uint64_t w0 = c * SCAN_CHUNK;
    uint64_t w1 = w0 + SCAN_CHUNK;
    if (w1 > nwords) w1 = nwords;
    __shared__ unsigned long long wave_tot[TPB / 64];
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) s_base = chunk_off[c];
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (uint64_t g = w0; g < w1; g += TPB) {
        const uint64_t i = g + threadIdx.x;
        uint64_t v = (i < w1) ? a[i] : 0ull;
        if (i == nwords - 1 && (nbits & 63)) v &= (1ull << (nbits & 63)) - 1ull;
        const uint32_t cnt = (uint32_t)__popcll(v);
        const uint32_t incl = psk::wave_incl_scan(cnt, lane);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        unsigned long long base = s_base;
        for (int w = 0; w < wave; ++w) base += wave_tot[w];
        uint64_t pos = base + incl - cnt;
        while (v) {
            if (pos < cap) out[pos] = i * 64 + (uint64_t)(__ffsll((unsigned long long)v) - 1);
            ++pos;
            v &= v - 1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (int w = 0; w < TPB / 64; ++w) t += wave_tot[w];
            s_base += t;
        }
        __syncthreads();
    }
}

__global__ void k_copy_count(const unsigned long long* src, uint64_t* dst) { *dst = *src; }

__global__ void k_set_u64(uint64_t* dst, uint64_t v) { *dst = v; }

}  // namespace

}  // namespace ps

// Scratch memory of ps_bits_scan_noalloc: the per-chunk set-bit counts of the
// scanned array, one u64 per SCAN_CHUNK words, then their total.  Every device
// the library runs on holds its own copy (8 MiB) until the process ends.
constexpr uint64_t NOALLOC_MAX_CHUNKS = 1ull << 20;  // arrays of up to 2^38 bits
static __device__ __align__(256) unsigned long long scratch[NOALLOC_MAX_CHUNKS + 1];

namespace ps {
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned long long* g_scratch[MAX_DEVICES] = {};  // device address of `scratch`, per device
}  // namespace

int bits_scratch_init(int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(PS_ERR_UNSUPPORTED, "device %d >= %d", device, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return PS_OK;
    void* p = nullptr;
    PS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = (unsigned long long*)p;
    return PS_OK;
}

namespace {
// ScanBits of nwords words into d_out with the chunk counts in tmp
// (nchunks + 1 entries, all written here); the total goes to *d_count.
int launch_scan(const uint64_t* d_a, uint64_t nwords, uint64_t nbits, unsigned long long* tmp, uint64_t nchunks,
                uint64_t* d_out, uint64_t cap, uint64_t* d_count, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_counts, dim3((uint32_t)nchunks), dim3(TPB), 0, st, d_a, nwords, nbits, tmp);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, tmp, nchunks);
    hipLaunchKernelGGL(k_scan_write, dim3((uint32_t)nchunks), dim3(TPB), 0, st, d_a, nwords, nbits, tmp, d_out, cap);
    hipLaunchKernelGGL(k_copy_count, dim3(1), dim3(1), 0, st, tmp + nchunks, d_count);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}
}  // namespace
}  // namespace ps

using namespace ps;

extern "C" int ps_bits_andnot_count(uint64_t* d_a, const uint64_t* d_excl1, const uint64_t* d_excl2, uint64_t nwords,
                                    uint64_t* d_count, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (nwords == 0) return PS_OK;
    if (!d_a || !d_excl1 || !d_count) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipLaunchKernelGGL(k_andnot_count, dim3(grid_for(nwords, 4)), dim3(TPB), 0, stream_of(s), d_a, d_excl1, d_excl2,
                       nwords, (unsigned long long*)d_count);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

extern "C" int ps_bits_count(const uint64_t* d_a, uint64_t nwords, uint64_t* d_count, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (nwords == 0) return PS_OK;
    if (!d_a || !d_count) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipLaunchKernelGGL(k_count, dim3(grid_for(nwords, 4)), dim3(TPB), 0, stream_of(s), d_a, nwords,
                       (unsigned long long*)d_count);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

extern "C" int ps_bits_hash(const uint64_t* d_a, uint64_t nwords, uint64_t word_offset, uint64_t* d_hash, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (nwords == 0) return PS_OK;
    if (!d_a || !d_hash) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipLaunchKernelGGL(k_hash, dim3(grid_for(nwords, 4)), dim3(TPB), 0, stream_of(s), d_a, nwords, word_offset,
                       (unsigned long long*)d_hash);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

extern "C" int ps_bits_set_list(uint64_t* d_a, uint64_t nbits, const uint64_t* d_idx, uint64_t count,
                                uint64_t* d_rejected, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (count == 0) return PS_OK;
    if (!d_a || !d_idx) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipLaunchKernelGGL(k_set_list, dim3(grid_for(count, 4)), dim3(TPB), 0, stream_of(s), (uint32_t*)d_a, nbits, d_idx,
                       count, (unsigned long long*)d_rejected);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

extern "C" int ps_bits_scan(const uint64_t* d_a, uint64_t nbits, uint64_t* d_out, uint64_t cap, uint64_t* d_count,
                            ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_count) return fail(PS_ERR_INVALID_ARGUMENT, "NULL count pointer");
    hipStream_t st = stream_of(s);
    const uint64_t nwords = (nbits + 63) / 64;
    if (nwords == 0) {
        PS_HIP_CHECK(hipMemsetAsync(d_count, 0, 8, st));
        return PS_OK;
    }
    if (!d_a || (cap && !d_out)) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    const uint64_t nchunks = (nwords + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nchunks > 0xFFFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "array of %llu bits too large",
                                            (unsigned long long)nbits);
    unsigned long long* tmp = nullptr;
    PS_HIP_CHECK(hipMallocAsync((void**)&tmp, (nchunks + 1) * 8, st));
    if (int r = launch_scan(d_a, nwords, nbits, tmp, nchunks, d_out, cap, d_count, st)) {
        (void)hipFreeAsync(tmp, st);
        return r;
    }
    uint64_t total = 0;
    PS_HIP_CHECK(hipMemcpyAsync(&total, tmp + nchunks, 8, hipMemcpyDeviceToHost, st));
    PS_HIP_CHECK(hipFreeAsync(tmp, st));
    PS_HIP_CHECK(hipStreamSynchronize(st));
    if (total > cap) return fail(PS_ERR_CAPACITY, "ps_bits_scan: %llu set bits exceed capacity %llu",
                                 (unsigned long long)total, (unsigned long long)cap);
    return PS_OK;
    PS_API_END
}

// The same scan with its chunk counts in the static `scratch` array of this
// device: no allocation and no wait, so a graph capture may hold it.  Clears
// nothing first: every scratch entry a call reads it has written itself.
extern "C" int ps_bits_scan_noalloc(const uint64_t* d_a, uint64_t nbits, uint64_t* d_out, uint64_t cap,
                                    uint64_t* d_count, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_count) return fail(PS_ERR_INVALID_ARGUMENT, "NULL count pointer");
    hipStream_t st = stream_of(s);
    const uint64_t nwords = (nbits + 63) / 64;
    const uint64_t nchunks = (nwords + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nchunks > NOALLOC_MAX_CHUNKS)
        return fail(PS_ERR_INVALID_ARGUMENT,
                    "ps_bits_scan_noalloc: %llu bits exceed the static scratch limit of 2^38 bits",
                    (unsigned long long)nbits);
    if (nwords == 0) {
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, d_count, 0ull);
        PS_HIP_CHECK(hipGetLastError());
        return PS_OK;
    }
    if (!d_a || (cap && !d_out)) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    int dev = -1;
    PS_HIP_CHECK(hipGetDevice(&dev));
    unsigned long long* tmp = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (dev >= 0 && dev < MAX_DEVICES) tmp = g_scratch[dev];
    }
    if (!tmp) return fail(PS_ERR_NOT_INITIALIZED, "ps_bits_scan_noalloc: device %d was not set up by ps_init", dev);
    return launch_scan(d_a, nwords, nbits, tmp, nchunks, d_out, cap, d_count, st);
    PS_API_END
}
