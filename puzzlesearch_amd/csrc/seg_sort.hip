// Device prefix sums and stable counting sorts (declarations and the method
// in seg_sort.h).
#include "seg_sort.h"

#include <algorithm>
#include <vector>

#include "devbuf.h"
#include "ps_internal.h"

namespace psk {

typedef unsigned long long ull;

constexpr int SC_TPB = 256;
constexpr int SC_PER = 16;
constexpr uint64_t SC_TILE = (uint64_t)SC_TPB * SC_PER;  // values per tile of the prefix sum

__device__ __forceinline__ ull wave_incl_scan64(ull v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const ull t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ ull wave_sum64(ull v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// per tile of SC_TILE values: their sum
template <class T>
__global__ __launch_bounds__(SC_TPB) void k_sc_tile_sums(const T* __restrict__ in, uint64_t n, ull* __restrict__ sums) {
    const uint64_t t0 = (uint64_t)blockIdx.x * SC_TILE;
    ull s = 0;
#pragma unroll
    for (int k = 0; k < SC_PER; ++k) {
        const uint64_t i = t0 + (uint64_t)k * SC_TPB + threadIdx.x;
        if (i < n) s += (ull)in[i];
    }
    __shared__ ull red[SC_TPB / 64];
    s = wave_sum64(s);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        ull t = 0;
        for (int w = 0; w < SC_TPB / 64; ++w) t += red[w];
        sums[blockIdx.x] = t;
    }
}

// exclusive prefix sum of v[0..n) in place by one block; v[n] = the total
__global__ __launch_bounds__(1024) void k_sc_single(ull* v, uint64_t n) {
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = threadIdx.x * per, e = (b + per < n) ? b + per : n;
    ull s = 0;
    for (uint64_t i = b; i < e; ++i) s += v[i];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __shared__ ull wtot[16];
    const ull incl = wave_incl_scan64(s, lane);
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    ull before = 0, total = 0;
    for (int w = 0; w < 16; ++w) {
        before += (w < wave) ? wtot[w] : 0ull;
        total += wtot[w];
    }
    ull run = before + incl - s;
    for (uint64_t i = b; i < e; ++i) {
        const ull t = v[i];
        v[i] = run;
        run += t;
    }
    if (threadIdx.x == 0) v[n] = total;
}

// tile k: thread t owns values t0 + t*SC_PER .. +SC_PER-1
template <class T>
__global__ __launch_bounds__(SC_TPB) void k_sc_tile_apply(const T* __restrict__ in, uint64_t n,
                                                        const ull* __restrict__ base, ull* __restrict__ out) {
    const uint64_t t0 = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_PER;
    ull v[SC_PER];
    ull s = 0;
#pragma unroll
    for (int k = 0; k < SC_PER; ++k) {
        const uint64_t i = t0 + k;
        v[k] = (i < n) ? (ull)in[i] : 0ull;
        s += v[k];
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __shared__ ull wtot[SC_TPB / 64];
    const ull incl = wave_incl_scan64(s, lane);
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    ull run = base[blockIdx.x] + incl - s;
    for (int w = 0; w < wave; ++w) run += wtot[w];
#pragma unroll
    for (int k = 0; k < SC_PER; ++k) {
        const uint64_t i = t0 + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = base[gridDim.x];
}

__global__ void k_sc_zero_total(ull* out) { out[0] = 0ull; }

// ------------------------------------------------------ stable counting sort

constexpr uint32_t BK_MAX_KEYS = 8192;       // keys per counting pass (LDS histogram / cursors)
constexpr uint64_t BK_HIST_MAX = 1ull << 22;  // histogram entries per pass (keys x blocks)
constexpr uint32_t BK_BAD = 0xFFFFFFFFu;

__global__ void k_bk_present(const ull* __restrict__ kv, uint64_t m, uint32_t K, uint8_t* __restrict__ present,
                             ull* __restrict__ dropped) {
    ull bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t k = (uint32_t)(kv[i] >> 32);
        if (k < K) present[k] = 1;
        else ++bad;
    }
    bad = wave_sum64(bad);
    if (bad && threadIdx.x % 64 == 0) atomicAdd(dropped, bad);
}

// dense[k] = rank of key k among the present keys (BK_BAD if absent);
// keys_out lists the present keys ascending, *nd their number.  One block.
__global__ __launch_bounds__(1024) void k_bk_dense(const uint8_t* __restrict__ present, uint32_t K,
                                                  uint32_t* __restrict__ dense, uint32_t* __restrict__ keys_out,
                                                  uint32_t* __restrict__ nd) {
    const uint32_t per = (K + 1023) / 1024;
    const uint32_t b = threadIdx.x * per, e = (b + per < K) ? b + per : K;
    uint32_t c = 0;
    for (uint32_t k = b; k < e; ++k) c += present[k] ? 1u : 0u;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    __shared__ ull wtot[16];
    const ull incl = wave_incl_scan64(c, lane);
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    ull before = 0, total = 0;
    for (int w = 0; w < 16; ++w) {
        before += (w < wave) ? wtot[w] : 0ull;
        total += wtot[w];
    }
    uint32_t r = (uint32_t)(before + incl - c);
    for (uint32_t k = b; k < e; ++k) {
        if (present[k]) {
            dense[k] = r;
            keys_out[r] = k;
            ++r;
        } else {
            dense[k] = BK_BAD;
        }
    }
    if (threadIdx.x == 0) *nd = (uint32_t)total;
}

// the key of a counting pass: the dense rank of the entry's key minus lo
struct KeyDense {
    const uint32_t* dense;
    uint32_t K, lo;
    __device__ __forceinline__ uint32_t operator()(ull kv) const {
        const uint32_t k = (uint32_t)(kv >> 32);
        return (k < K) ? dense[k] - lo : BK_BAD;  // below lo wraps above every pass
    }
};

// the key of a counting pass: one byte of the value
struct KeyDigit {
    int shift;
    __device__ __forceinline__ uint32_t operator()(ull kv) const { return ((uint32_t)kv >> shift) & 255u; }
};

// hist[k * nblk + b] = entries of block b's chunk with pass key k (< nk)
template <class KF>
__global__ __launch_bounds__(256) void k_bk_hist(const ull* __restrict__ kv, uint64_t m, uint32_t chunk, KF kf,
                                                uint32_t nk, uint32_t* __restrict__ hist, uint32_t nblk) {
    extern __shared__ uint32_t h[];
    for (uint32_t k = threadIdx.x; k < nk; k += 256) h[k] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t b1 = (b0 + chunk < m) ? b0 + chunk : m;
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += 256) {
        const uint32_t k = kf(kv[i]);
        if (k < nk) atomicAdd(&h[k], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nk; k += 256) hist[(uint64_t)k * nblk + blockIdx.x] = h[k];
}

// per pass key k: the entries of that key, from the scanned histogram
__global__ void k_bk_counts(const ull* __restrict__ off, uint32_t nk, uint32_t nblk, ull* __restrict__ counts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) counts[k] = off[(uint64_t)(k + 1) * nblk] - off[(uint64_t)k * nblk];
}

// Stable scatter by one wave per block: the block's chunk is walked 64
// entries at a time in input order; lanes with equal pass keys are found with
// one ballot per key bit, ranked by lane, and written after the key's cursor,
// which the lowest such lane then advances.  A cursor is read and written only
// by this wave, in program order, so entries of one key keep their order.
template <class KF, bool OUT_KV>
__global__ __launch_bounds__(64) void k_bk_scatter(const ull* __restrict__ kv, uint64_t m, uint32_t chunk, KF kf,
                                                  uint32_t nk, int key_bits, const ull* __restrict__ off,
                                                  uint32_t nblk, ull* __restrict__ out_kv,
                                                  uint32_t* __restrict__ out_val) {
    extern __shared__ uint32_t cur[];
    const int lane = threadIdx.x;
    for (uint32_t k = lane; k < nk; k += 64) cur[k] = (uint32_t)off[(uint64_t)k * nblk + blockIdx.x];
    wave_fence();
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t b1 = (b0 + chunk < m) ? b0 + chunk : m;
    const ull lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t g = b0; g < b1; g += 64) {
        const uint64_t i = g + lane;
        const bool valid = i < b1;
        const ull v = valid ? kv[i] : 0ull;
        const uint32_t k = valid ? kf(v) : BK_BAD;
        const bool in = k < nk;
        ull peers = __ballot(in);
        for (int bit = 0; bit < key_bits; ++bit) {
            const bool one = (k >> bit) & 1u;
            const ull bm = __ballot(one);
            peers &= one ? bm : ~bm;
        }
        uint32_t pos = 0;
        if (in) {
            pos = cur[k] + (uint32_t)__popcll(peers & lt);
            if (OUT_KV) out_kv[pos] = v;
            else out_val[pos] = (uint32_t)v;
        }
        wave_fence();  // every lane has read its cursor before the leaders advance them
        if (in && (peers & lt) == 0ull) cur[k] += (uint32_t)__popcll(peers);
        wave_fence();
    }
}

}  // namespace psk

namespace ps {

using psk::ull;

namespace {

template <class T>
int scan_excl(const T* d_in, uint64_t n, ull* d_out, hipStream_t s) {
    if (n == 0) {
        hipLaunchKernelGGL(psk::k_sc_zero_total, dim3(1), dim3(1), 0, s, d_out);
        PS_HIP_CHECK(hipGetLastError());
        return PS_OK;
    }
    const uint64_t ntiles = (n + psk::SC_TILE - 1) / psk::SC_TILE;
    if (ntiles > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "prefix sum of %llu values", (ull)n);
    DevBuf sums;
    if (int r = sums.alloc((ntiles + 1) * 8, s)) return r;
    hipLaunchKernelGGL(psk::k_sc_tile_sums<T>, dim3((uint32_t)ntiles), dim3(psk::SC_TPB), 0, s, d_in, n, sums.as<ull>());
    hipLaunchKernelGGL(psk::k_sc_single, dim3(1), dim3(1024), 0, s, sums.as<ull>(), ntiles);
    hipLaunchKernelGGL(psk::k_sc_tile_apply<T>, dim3((uint32_t)ntiles), dim3(psk::SC_TPB), 0, s, d_in, n,
                       (const ull*)sums.as<ull>(), d_out);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

int key_bits_for(uint32_t nk) {
    int b = 1;
    while ((1u << b) < nk) ++b;
    return b;
}

// chunk of entries per block: a multiple of 64 that keeps nk x blocks within
// the histogram bound
uint32_t chunk_for(uint64_t m, uint32_t nk) {
    uint64_t c = 4096;
    while ((m + c - 1) / c * (uint64_t)nk > psk::BK_HIST_MAX) c *= 2;
    return (uint32_t)c;
}

// One stable counting pass with pass keys kf(kv) in [0, nk): writes to out_kv
// (or the values to out_val), at out + 0; returns the per-key sizes (device
// counts[nk]) when counts is given.
template <class KF, bool OUT_KV>
int counting_pass(const ull* d_kv, uint64_t m, KF kf, uint32_t nk, ull* out_kv, uint32_t* out_val, ull* d_counts,
                  hipStream_t s) {
    const uint32_t chunk = chunk_for(m, nk);
    const uint64_t nblk64 = (m + chunk - 1) / chunk;
    if (nblk64 == 0) return PS_OK;
    const uint32_t nblk = (uint32_t)nblk64;
    const uint64_t nh = (uint64_t)nk * nblk;
    DevBuf hist, off;
    if (int r = hist.alloc(nh * 4, s)) return r;
    if (int r = off.alloc((nh + 1) * 8, s)) return r;
    hipLaunchKernelGGL((psk::k_bk_hist<KF>), dim3(nblk), dim3(256), nk * 4, s, d_kv, m, chunk, kf, nk,
                       hist.as<uint32_t>(), nblk);
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl<uint32_t>(hist.as<uint32_t>(), nh, off.as<ull>(), s)) return r;
    if (d_counts) {
        hipLaunchKernelGGL(psk::k_bk_counts, dim3((nk + 255) / 256), dim3(256), 0, s, (const ull*)off.as<ull>(), nk,
                           nblk, d_counts);
        PS_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL((psk::k_bk_scatter<KF, OUT_KV>), dim3(nblk), dim3(64), nk * 4, s, d_kv, m, chunk, kf, nk,
                       key_bits_for(nk), (const ull*)off.as<ull>(), nblk, out_kv, out_val);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

}  // namespace

int scan_excl_u32(const uint32_t* d_in, uint64_t n, ull* d_out, hipStream_t s) { return scan_excl(d_in, n, d_out, s); }
int scan_excl_u64(const ull* d_in, uint64_t n, ull* d_out, hipStream_t s) { return scan_excl(d_in, n, d_out, s); }

int stable_group_by_key(const ull* d_kv, uint64_t m, uint32_t key_space, uint32_t* d_values_out,
                        std::vector<uint32_t>& keys, std::vector<uint64_t>& counts, uint64_t* dropped,
                        hipStream_t s) {
    keys.clear();
    counts.clear();
    if (dropped) *dropped = 0;
    if (m == 0) return PS_OK;
    if (m > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "at most 2^31-1 entries per grouping");
    if (key_space == 0 || key_space > (1u << 26)) return fail(PS_ERR_INVALID_ARGUMENT, "key space %u", key_space);
    DevBuf present, dense, keys_d, small;
    if (int r = present.alloc(key_space, s)) return r;
    if (int r = dense.alloc((uint64_t)key_space * 4, s)) return r;
    if (int r = keys_d.alloc((uint64_t)key_space * 4, s)) return r;
    if (int r = small.alloc(16, s)) return r;
    if (int r = launch_fill(present.p, 0, key_space, s)) return r;
    if (int r = launch_fill(small.p, 0, 16, s)) return r;
    hipLaunchKernelGGL(psk::k_bk_present, dim3((uint32_t)std::min<uint64_t>(blocks_for(m, 256), 8192)), dim3(256), 0, s,
                       d_kv, m, key_space, present.as<uint8_t>(), small.as<ull>());
    hipLaunchKernelGGL(psk::k_bk_dense, dim3(1), dim3(1024), 0, s, (const uint8_t*)present.as<uint8_t>(), key_space,
                       dense.as<uint32_t>(), keys_d.as<uint32_t>(), (uint32_t*)(small.as<ull>() + 1));
    PS_HIP_CHECK(hipGetLastError());
    ull h[2] = {0, 0};
    PS_HIP_CHECK(hipMemcpyAsync(h, small.p, 16, hipMemcpyDeviceToHost, s));
    if (int r = stream_wait(s)) return r;
    const uint32_t nd = (uint32_t)h[1];
    if (dropped) *dropped = h[0];
    if (nd == 0) return PS_OK;
    if (int r = to_host(keys, keys_d.p, nd, s)) return r;
    DevBuf cnt;
    if (int r = cnt.alloc((uint64_t)nd * 8, s)) return r;
    // passes over ranges of at most BK_MAX_KEYS dense keys; each pass writes
    // after the entries of the keys before it
    uint64_t base = 0;
    std::vector<uint64_t> pc;
    for (uint32_t lo = 0; lo < nd; lo += psk::BK_MAX_KEYS) {
        const uint32_t nk = std::min<uint32_t>(psk::BK_MAX_KEYS, nd - lo);
        psk::KeyDense kf{dense.as<uint32_t>(), key_space, lo};
        if (int r = counting_pass<psk::KeyDense, false>(d_kv, m, kf, nk, nullptr, d_values_out + base,
                                                        cnt.as<ull>() + lo, s))
            return r;
        if (int r = to_host(pc, cnt.as<ull>() + lo, nk, s)) return r;
        for (uint64_t c : pc) {
            counts.push_back(c);
            base += c;
        }
    }
    if (base + h[0] != m)
        return fail(PS_ERR_INTERNAL, "grouping placed %llu of %llu entries", (ull)base, (ull)(m - h[0]));
    return PS_OK;
}

int stable_sort_by_value(ull* d_kv, ull* d_tmp, uint64_t m, int value_bits, hipStream_t s) {
    if (m == 0) return PS_OK;
    if (m > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "at most 2^31-1 entries per sort");
    ull* src = d_kv;
    ull* dst = d_tmp;
    for (int shift = 0; shift < value_bits; shift += 8) {
        if (int r = counting_pass<psk::KeyDigit, true>(src, m, psk::KeyDigit{shift}, 256u, dst, nullptr, nullptr, s))
            return r;
        std::swap(src, dst);
    }
    if (src != d_kv) PS_HIP_CHECK(hipMemcpyAsync(d_kv, src, m * 8, hipMemcpyDeviceToDevice, s));
    return PS_OK;
}

}  // namespace ps
