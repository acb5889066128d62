// Frontier codecs on the GPU, byte-identical to the reference CPU codecs:
//   StreamVInt (Common/StreamVInt.cpp:149-300)
//   OpBitsCompression<2/4/8> (Common/OpBitsCompression.cpp:10-76)
//   FrontierCompression bitmap-or-StreamVInt chunks (Common/FrontierCompression.cpp:28-150)
//   bucketing of cross-segment children by (op, target segment), the grouping
//   SmartMultiplexor::Add performs (Common/Multiplexor.cpp:200-212), in
//   canonical (sorted) order.
//
// Values come in chunks (chunk c = values[starts[c], starts[c+1])), each encoded
// independently and concatenated, as a writer appends encoded chunks to a
// segment file.  Encoding is parallel per 4-value tuple: tuple sizes, a scan
// per chunk for byte offsets, then the writes.  Decoding first walks each
// chunk's tuple descriptors (one thread per chunk: a descriptor's position
// depends on every earlier tuple's size), then decodes tuples in parallel and
// undoes the delta coding with a scan per chunk.
#include <algorithm>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "codec_internal.h"
#include "devbuf.h"
#include "ps_internal.h"
#include "seg_sort.h"

namespace ps {
namespace {

constexpr uint32_t CODEC_MAP = 0x80000000u;

__device__ __forceinline__ uint32_t vlen(uint32_t v) { return (v >> 24) ? 4u : (v >> 16) ? 3u : (v >> 8) ? 2u : 1u; }

// chunk of global tuple t (tuple_base has nchunks+1 entries)
__device__ __forceinline__ uint32_t chunk_of(const uint64_t* tuple_base, uint32_t nchunks, uint64_t t) {
    uint32_t lo = 0, hi = nchunks;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tuple_base[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// inclusive sum over the 64 lanes of a wave
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

struct SviPlan {
    // per chunk
    const uint64_t* starts;      // value starts, nchunks+1
    const uint64_t* tuple_base;  // nchunks+1
    const uint32_t* counts;
    uint32_t nchunks;
    uint64_t ntuples;
    uint32_t max_tuples;  // of one chunk
    int32_t grid2d;       // launches are (tuple blocks of a chunk) x (chunk): no search for the chunk
};

// The tuple a thread encodes and its chunk; false for a thread past the end.
__device__ __forceinline__ bool svi_tuple(const SviPlan& pl, uint64_t& t, uint32_t& c) {
    if (pl.grid2d) {
        c = blockIdx.y;
        t = pl.tuple_base[c] + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        return t < pl.tuple_base[c + 1];
    }
    t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pl.ntuples) return false;
    c = chunk_of(pl.tuple_base, pl.nchunks, t);
    return true;
}

// launch shape of the per-tuple kernels for a plan
inline dim3 svi_grid(const SviPlan& pl) {
    if (pl.grid2d) return dim3((pl.max_tuples + 255) / 256, pl.nchunks);
    return dim3((uint32_t)((pl.ntuples + 255) / 256));
}

// delta of value j of chunk c (padding values are 0)
__device__ __forceinline__ uint32_t delta_at(const uint32_t* v, uint64_t start, uint32_t count, uint32_t j) {
    if (j >= count) return 0u;
    const uint32_t x = v[start + j];
    return j ? x - v[start + j - 1] : x;
}

__global__ void k_svi_tuple_sizes(const uint32_t* v, SviPlan pl, uint32_t* sizes) {
    uint64_t t;
    uint32_t c;
    if (!svi_tuple(pl, t, c)) return;
    const uint32_t lt = (uint32_t)(t - pl.tuple_base[c]);
    uint32_t sz = 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) sz += vlen(delta_at(v, pl.starts[c], pl.counts[c], lt * 4 + j));
    sizes[t] = sz;
}

// exclusive scan of tuple sizes within each chunk (one block per chunk);
// chunk_payload[c] = sum of the chunk's tuple sizes.  The chunk is scanned
// in tiles of 4*TPB sizes, four consecutive sizes per thread (a wave reads
// 1 KB at a time), the running total carried from tile to tile.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_chunk_scan_u32(uint32_t* vals, const uint64_t* base, uint64_t* chunk_total) {
    constexpr int NW = TPB / 64;
    const uint32_t c = blockIdx.x;
    const uint64_t b = base[c], e = base[c + 1];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t s_w[2][NW];
    uint32_t carry = 0, par = 0;
    for (uint64_t q0 = b; q0 < e; q0 += 4 * TPB, par ^= 1u) {
        const uint64_t i = q0 + 4 * (uint64_t)threadIdx.x;
        uint32_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = i + k < e ? vals[i + k] : 0u;
        const uint32_t s4 = x[0] + x[1] + x[2] + x[3];
        const uint32_t incl = wave_incl_scan_u32(s4, lane);
        if (lane == 63) s_w[par][wave] = incl;
        __syncthreads();
        // s_w[par] is written again two tiles later, after the next tile's
        // barrier, which every thread reaches only once done here
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t t = s_w[par][w];
            before += (w < wave) ? t : 0u;
            total += t;
        }
        uint32_t run = carry + before + incl - s4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i + k < e) vals[i + k] = run;
            run += x[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) chunk_total[c] = carry;
}

__global__ void k_svi_write(const uint32_t* v, const uint8_t* bounds, SviPlan pl, const uint32_t* tuple_off,
                            const uint64_t* chunk_out, const uint64_t* chunk_payload, uint8_t* out) {
    uint64_t t;
    uint32_t c;
    if (!svi_tuple(pl, t, c)) return;
    const uint32_t lt = (uint32_t)(t - pl.tuple_base[c]);
    const uint64_t start = pl.starts[c];
    const uint32_t count = pl.counts[c];
    uint8_t* o = out + chunk_out[c];
    if (lt == 0) {
        o[0] = (uint8_t)count;
        o[1] = (uint8_t)(count >> 8);
        o[2] = (uint8_t)(count >> 16);
        o[3] = (uint8_t)(count >> 24);
    }
    uint8_t* p = o + 4 + tuple_off[t];
    uint8_t desc = 0;
    int pos = 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t d = delta_at(v, start, count, lt * 4 + j);
        const uint32_t len = vlen(d);
        desc |= (uint8_t)((len - 1) << (2 * j));
        for (uint32_t b = 0; b < len; ++b) p[pos++] = (uint8_t)(d >> (8 * b));
    }
    p[0] = desc;
    if (bounds && (lt & 1) == 0) {
        // eight 4-bit bounds per 4 bytes, folded like _EncodeBoundsTuple
        uint64_t all = 0;
        for (int j = 0; j < 8; ++j) {
            const uint32_t idx = lt * 4 + j;
            const uint64_t bv = idx < count ? bounds[start + idx] : 0;
            all |= bv << (8 * j);
        }
        const uint32_t w = (uint32_t)(all | (all >> 28));
        uint8_t* q = o + 4 + chunk_payload[c] + (lt / 2) * 4;
        q[0] = (uint8_t)w;
        q[1] = (uint8_t)(w >> 8);
        q[2] = (uint8_t)(w >> 16);
        q[3] = (uint8_t)(w >> 24);
    }
}

// ---------------------------------------------------------------- decode

__device__ __forceinline__ uint32_t rd32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

__global__ void k_read_counts(const uint8_t* in, const uint64_t* in_off, uint32_t nchunks, uint32_t* counts,
                              uint32_t* is_bitmap) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    if (in_off[c + 1] == in_off[c]) {  // an empty chunk encodes to no bytes
        counts[c] = 0;
        if (is_bitmap) is_bitmap[c] = 0;
        return;
    }
    const uint8_t* p = in + in_off[c];
    const uint32_t h = rd32(p);
    if (is_bitmap && (h & CODEC_MAP)) {
        is_bitmap[c] = 1;
        counts[c] = h & ~CODEC_MAP;  // bitmap words; the value count comes later
    } else {
        if (is_bitmap) is_bitmap[c] = 0;
        counts[c] = h;
    }
}

// Descriptor walk of one chunk per block, in parallel: a chunk's tuples
// (descriptor byte + 1..4 value bytes each, 5..17 bytes) tile its payload
// [4, E), E = chunk bytes - bound bytes.  Thread t owns payload bytes
// [s0, s1); the first tuple starting at or after s0 starts at s0 + o,
// o in [0, 16].  Each thread walks its range from all 17 entry offsets at once
// (independent loads), recording where each walk leaves the range and how many
// tuples it passed; thread 0 chains the ranges from entry 0; then every thread
// walks its range once more from its true entry, writing the tuple offsets.
// A chunk whose tuples do not end exactly at E counts in *err.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_svi_walk_par(const uint8_t* in, const uint64_t* in_off,
                                                      const uint64_t* tuple_base, const uint32_t* select,
                                                      int with_bounds, uint32_t* tuple_off, uint32_t* payload_end,
                                                      uint32_t* err) {
    constexpr int NE = 17;
    const uint32_t k = blockIdx.x;
    const uint32_t c = select ? select[k] : k;
    const uint8_t* p = in + in_off[c];
    const uint64_t clen = in_off[c + 1] - in_off[c];
    const uint64_t T = tuple_base[k + 1] - tuple_base[k];
    const uint64_t bnd = with_bounds ? T * 2 : 0;  // (T*4 padded values) / 2
    const uint64_t E = clen >= 4 + bnd ? clen - bnd : 4;
    // T tuples take 5T..17T bytes: any other payload length is refused
    // before it is walked (this also bounds the per-range tuple counts below)
    if (E - 4 < 5 * T || E - 4 > 17 * T) {
        if (threadIdx.x == 0) atomicAdd(err, 1u);
        return;
    }
    const uint64_t L = (E - 4 + TPB - 1) / TPB;
    const uint64_t s0 = 4 + (uint64_t)threadIdx.x * L;
    const uint64_t s1 = s0 + L < E ? s0 + L : (s0 < E ? E : s0);
    __shared__ uint8_t s_ex[TPB][NE];
    __shared__ uint16_t s_n[TPB][NE];  // tuples of a range: at most its bytes / 5
    __shared__ uint32_t s_entry[TPB], s_base[TPB];
    {
        uint64_t pos[NE];
        uint32_t n[NE];
#pragma unroll
        for (int o = 0; o < NE; ++o) {
            pos[o] = s0 + (uint64_t)o;
            n[o] = 0;
        }
        // all 17 descriptor loads of a step are issued before any is used; a
        // walk that has left the range reads byte 0 of the chunk instead
        // (a selected chunk is not empty) and does not move
        bool any = s0 < s1;
        while (any) {
            uint32_t d[NE];
#pragma unroll
            for (int o = 0; o < NE; ++o) d[o] = p[pos[o] < s1 ? pos[o] : 0];
            any = false;
#pragma unroll
            for (int o = 0; o < NE; ++o) {
                const bool live = pos[o] < s1;
                const uint32_t x = d[o];
                pos[o] += live ? 5 + (x & 3) + ((x >> 2) & 3) + ((x >> 4) & 3) + ((x >> 6) & 3) : 0;
                n[o] += live ? 1u : 0u;
                any |= live;
            }
        }
#pragma unroll
        for (int o = 0; o < NE; ++o) {
            s_ex[threadIdx.x][o] = (uint8_t)(pos[o] - s1);  // 0..16
            s_n[threadIdx.x][o] = (uint16_t)n[o];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t o = 0;
        uint64_t base = 0;
        for (int t = 0; t < TPB; ++t) {
            s_entry[t] = o;
            s_base[t] = (uint32_t)base;
            base += s_n[t][o];
            o = s_ex[t][o];
        }
        payload_end[k] = (uint32_t)E;
        if (o != 0 || base != T || clen < 4 + bnd) atomicAdd(err, 1u);
    }
    __syncthreads();
    uint64_t pos = s0 + s_entry[threadIdx.x];
    uint64_t i = s_base[threadIdx.x];
    uint32_t* toff = tuple_off + tuple_base[k];
    while (pos < s1 && i < T) {
        toff[i++] = (uint32_t)pos;
        const uint32_t d = p[pos];
        pos += 5 + (d & 3) + ((d >> 2) & 3) + ((d >> 4) & 3) + ((d >> 6) & 3);
    }
}

// grid2d: block row y decodes selected chunk y (no search for a tuple's chunk)
__global__ void k_svi_decode_tuples(const uint8_t* in, const uint64_t* in_off, const uint64_t* tuple_base,
                                    const uint32_t* select, uint32_t nsel, uint64_t ntuples, int grid2d,
                                    const uint32_t* tuple_off, uint32_t* deltas) {
    uint64_t t;
    uint32_t k;
    if (grid2d) {
        k = blockIdx.y;
        t = tuple_base[k] + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= tuple_base[k + 1]) return;
    } else {
        t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= ntuples) return;
        k = chunk_of(tuple_base, nsel, t);
    }
    const uint32_t c = select ? select[k] : k;
    const uint8_t* p = in + in_off[c] + tuple_off[t];
    const uint32_t desc = p[0];
    int pos = 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t len = 1 + ((desc >> (2 * j)) & 3);
        uint32_t v = 0;
        for (uint32_t b = 0; b < len; ++b) v |= (uint32_t)p[pos++] << (8 * b);
        deltas[t * 4 + j] = v;
    }
}

// inclusive scan of one chunk's deltas (one block per chunk), first `count`
// results to the output; bounds decoded alongside.  The chunk is scanned in
// tiles of 4*TPB deltas: each thread loads four consecutive deltas as one
// 16-byte vector (coalesced), the block scans the per-thread sums, and the
// running total carries from tile to tile.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_svi_finish(const uint32_t* deltas, const uint64_t* tuple_base,
                                                    const uint32_t* counts, const uint32_t* select,
                                                    const uint64_t* value_starts, uint32_t* out,
                                                    const uint8_t* in, const uint64_t* in_off,
                                                    const uint32_t* payload_end, uint8_t* bounds_out) {
    constexpr int NW = TPB / 64;
    const uint32_t k = blockIdx.x;
    const uint32_t c = select ? select[k] : k;
    const uint64_t t0 = tuple_base[k], t1 = tuple_base[k + 1];  // one tuple = four deltas, 16-byte aligned
    const uint32_t count = counts[c];
    uint32_t* o = out + value_starts[c];
    const uint4* d4 = reinterpret_cast<const uint4*>(deltas);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t s_w[2][NW];
    uint32_t carry = 0, par = 0;
    for (uint64_t q0 = t0; q0 < t1; q0 += TPB, par ^= 1u) {
        const uint64_t q = q0 + threadIdx.x;
        uint4 v = q < t1 ? d4[q] : make_uint4(0u, 0u, 0u, 0u);
        v.y += v.x;
        v.z += v.y;
        v.w += v.z;
        const uint32_t incl = wave_incl_scan_u32(v.w, lane);
        if (lane == 63) s_w[par][wave] = incl;
        __syncthreads();
        // s_w[par] is written again two tiles later, after the next tile's
        // barrier, which every thread reaches only once done here
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t x = s_w[par][w];
            before += (w < wave) ? x : 0u;
            total += x;
        }
        const uint32_t base = carry + before + incl - v.w;
        const uint64_t j = (q - t0) * 4;
        if (j + 3 < count) {
            o[j] = base + v.x;
            o[j + 1] = base + v.y;
            o[j + 2] = base + v.z;
            o[j + 3] = base + v.w;
        } else if (q < t1) {
            const uint32_t r[4] = {v.x, v.y, v.z, v.w};
            for (int e = 0; e < 4; ++e)
                if (j + e < count) o[j + e] = base + r[e];
        }
        carry += total;
    }
    if (bounds_out) {
        const uint8_t* q = in + in_off[c] + payload_end[k];
        uint8_t* bo = bounds_out + value_starts[c];
        for (uint64_t g = threadIdx.x; g * 8 < count; g += TPB) {
            const uint64_t stored = rd32(q + g * 4);
            const uint64_t all = (stored & 0x0F0F0F0Full) | ((stored & 0xF0F0F0F0ull) << 28);
            for (int j = 0; j < 8; ++j) {
                const uint64_t idx = g * 8 + j;
                if (idx < count) bo[idx] = (uint8_t)(all >> (8 * j));
            }
        }
    }
}

// ---------------------------------------------------------------- bitmap chunks

// per chunk: first, last, whether every value v (index >= 1) lies in [first+1, last]
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fc_inspect(const uint32_t* v, const uint64_t* starts, uint32_t* first,
                                                    uint32_t* last, uint32_t* ok) {
    const uint32_t c = blockIdx.x;
    const uint64_t b = starts[c], e = starts[c + 1];
    if (b == e) {
        if (threadIdx.x == 0) ok[c] = 0;
        return;
    }
    const uint32_t f = v[b], l = v[e - 1];
    int good = l > f;
    // four independent loads per thread and step
    for (uint64_t i = b + 1 + threadIdx.x; i < e && good; i += 4 * TPB) {
        uint32_t x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = i + k * TPB < e ? v[i + k * TPB] : l;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (!(x[k] > f && x[k] <= l)) good = 0;
    }
    good = __syncthreads_and(good);
    if (threadIdx.x == 0) {
        first[c] = f;
        last[c] = l;
        ok[c] = good;
    }
}

__global__ void k_fc_bitmap_set(const uint32_t* v, const uint64_t* starts, const uint32_t* select, uint32_t nsel,
                                const uint64_t* word_base, const uint32_t* first, unsigned long long* words) {
    // one block per selected chunk
    const uint32_t k = blockIdx.x;
    if (k >= nsel) return;
    const uint32_t c = select[k];
    const uint64_t b = starts[c], e = starts[c + 1];
    const uint32_t f = first[c];
    unsigned long long* w = words + word_base[k];
    // each thread ORs a run of RUN consecutive values, one atomic per word the
    // run touches (sorted values put neighbours in the same word)
    constexpr uint64_t RUN = 32;
    for (uint64_t r0 = b + 1 + (uint64_t)threadIdx.x * RUN; r0 < e; r0 += (uint64_t)blockDim.x * RUN) {
        const uint64_t r1 = r0 + RUN < e ? r0 + RUN : e;
        uint64_t cw = ~0ull;
        unsigned long long m = 0;
        for (uint64_t i = r0; i < r1; ++i) {
            const uint64_t val = (uint64_t)v[i] - f - 1;
            if ((val >> 6) != cw) {
                if (m) atomicOr(&w[cw], m);
                cw = val >> 6;
                m = 0;
            }
            m |= 1ull << (val & 63);
        }
        if (m) atomicOr(&w[cw], m);
    }
}

__global__ void k_fc_bitmap_write(const uint32_t* select, uint32_t nsel, const uint64_t* word_base,
                                  const uint32_t* first, const unsigned long long* words, const uint64_t* out_off,
                                  uint8_t* out) {
    const uint32_t k = blockIdx.x;
    if (k >= nsel) return;
    const uint32_t c = select[k];
    const uint64_t nw = word_base[k + 1] - word_base[k];
    uint8_t* o = out + out_off[c];
    if (threadIdx.x == 0) {
        const uint32_t h = (uint32_t)nw | CODEC_MAP, f = first[c];
        for (int j = 0; j < 4; ++j) {
            o[j] = (uint8_t)(h >> (8 * j));
            o[4 + j] = (uint8_t)(f >> (8 * j));
        }
    }
    const uint8_t* src = reinterpret_cast<const uint8_t*>(words + word_base[k]);
    for (uint64_t i = threadIdx.x; i < nw * 8; i += blockDim.x) o[8 + i] = src[i];
}

__global__ void k_fc_bitmap_count(const uint8_t* in, const uint64_t* in_off, const uint32_t* select, uint32_t nsel,
                                  const uint32_t* words, uint32_t* counts) {
    const uint32_t k = blockIdx.x;
    if (k >= nsel) return;
    const uint32_t c = select[k];
    const uint8_t* p = in + in_off[c] + 8;
    uint32_t s = 0;
    for (uint64_t i = threadIdx.x; i < (uint64_t)words[c] * 8; i += blockDim.x) s += __popc(p[i]);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ uint32_t part[16];
    if (threadIdx.x % 64 == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += part[w];
        counts[c] = 1 + t;
    }
}

__global__ void k_fc_bitmap_decode(const uint8_t* in, const uint64_t* in_off, const uint32_t* select, uint32_t nsel,
                                   const uint32_t* words, const uint64_t* value_starts, uint32_t* out) {
    // one wave per selected chunk: ordered expansion of the bitmap
    const uint32_t k = blockIdx.x;
    if (k >= nsel) return;
    const uint32_t c = select[k];
    const uint8_t* p = in + in_off[c];
    const uint32_t f = rd32(p + 4);
    uint32_t* o = out + value_starts[c];
    const int lane = threadIdx.x;
    if (lane == 0) o[0] = f;
    uint64_t written = 1;
    const uint64_t nw = words[c];
    for (uint64_t g = 0; g < nw; g += 64) {
        const uint64_t wi = g + lane;
        uint64_t w = 0;
        if (wi < nw)
            for (int j = 0; j < 8; ++j) w |= (uint64_t)p[8 + wi * 8 + j] << (8 * j);
        const uint32_t cnt = (uint32_t)__popcll(w);
        uint32_t incl = cnt;
        for (int o2 = 1; o2 < 64; o2 <<= 1) {
            const uint32_t x = __shfl_up(incl, o2, 64);
            if (lane >= o2) incl += x;
        }
        uint64_t pos = written + incl - cnt;
        while (w) {
            o[pos++] = f + (uint32_t)(wi * 64 + (uint64_t)(__ffsll((unsigned long long)w) - 1)) + 1;
            w &= w - 1;
        }
        written += __shfl(incl, 63, 64);
    }
}

// ---------------------------------------------------------------- op bits

__global__ void k_opbits_encode(int bits, const uint8_t* ops, uint64_t groups, uint8_t* out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    uint64_t all = 0;
    for (int j = 0; j < 8; ++j) all |= (uint64_t)ops[g * 8 + j] << (8 * j);
    if (bits == 2) {
        const uint16_t v = (uint16_t)(all | (all >> 14) | (all >> 28) | (all >> 42));
        out[g * 2] = (uint8_t)v;
        out[g * 2 + 1] = (uint8_t)(v >> 8);
    } else if (bits == 4) {
        const uint32_t v = (uint32_t)(all | (all >> 28));
        for (int j = 0; j < 4; ++j) out[g * 4 + j] = (uint8_t)(v >> (8 * j));
    } else {
        for (int j = 0; j < 8; ++j) out[g * 8 + j] = (uint8_t)(all >> (8 * j));
    }
}

__global__ void k_opbits_decode(int bits, const uint8_t* in, uint64_t groups, uint8_t* ops) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= groups) return;
    uint64_t out;
    if (bits == 2) {
        const uint64_t s = (uint64_t)in[g * 2] | ((uint64_t)in[g * 2 + 1] << 8);
        out = (s & 0x0303ull) | ((s & 0x0C0Cull) << 14) | ((s & 0x3030ull) << 28) | ((s & 0xC0C0ull) << 42);
    } else if (bits == 4) {
        const uint64_t s = rd32(in + g * 4);
        out = (s & 0x0F0F0F0Full) | ((s & 0xF0F0F0F0ull) << 28);
    } else {
        out = 0;
        for (int j = 0; j < 8; ++j) out |= (uint64_t)in[g * 8 + j] << (8 * j);
    }
    for (int j = 0; j < 8; ++j) ops[g * 8 + j] = (uint8_t)(out >> (8 * j));
}

// ---------------------------------------------------------------- buckets

// (op * nseg + segment) << 32 | in-segment index; a child outside the
// segments or an op >= nops gets the key 0xFFFFFFFF, which the grouping drops
// and counts
__global__ void k_bucket_keys(const uint64_t* children, const uint8_t* ops, uint64_t n, int seg_bits, uint32_t nseg,
                              uint32_t nops, unsigned long long* kv) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = children[i];
    const uint64_t seg = c >> seg_bits;
    const uint32_t op = ops[i];
    const uint64_t key = (op < nops && seg < nseg) ? (uint64_t)op * nseg + seg : 0xFFFFFFFFull;
    kv[i] = (key << 32) | (c & ((1ull << seg_bits) - 1));
}


// ------------------------------------------------------------ host helpers

// SVI-encodes the chunks in `sel` (all chunks when empty) of values[starts];
// chunk k's bytes go to out + out_off[sel[k]] where out_off is computed here
// unless `fixed_out_off` is given.
struct SviEncodeResult {
    std::vector<uint64_t> bytes;  // per selected chunk
};

int svi_sizes(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
              const std::vector<uint32_t>& sel, hipStream_t s, DevBuf& d_starts_sel, DevBuf& d_tbase, DevBuf& d_counts,
              DevBuf& d_toff, DevBuf& d_payload, std::vector<uint64_t>& chunk_bytes, SviPlan& pl) {
    const uint32_t n = (uint32_t)sel.size();
    std::vector<uint64_t> st(n + 1), tb(n + 1);
    std::vector<uint32_t> cnt(n);
    uint64_t t = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t c0 = starts[sel[k]], c1 = starts[sel[k] + 1];
        if (c1 - c0 > MAX_CHUNK_VALUES) return fail(PS_ERR_INVALID_ARGUMENT, "chunk of %llu values exceeds %u",
                                                    (unsigned long long)(c1 - c0), MAX_CHUNK_VALUES);
        st[k] = c0;
        cnt[k] = (uint32_t)(c1 - c0);
        tb[k] = t;
        t += (uint64_t)((cnt[k] + 7) / 8 * 8) / 4;
    }
    st[n] = n ? starts[sel[n - 1] + 1] : 0;
    tb[n] = t;
    if (int r = to_device(d_starts_sel, st, s)) return r;
    if (int r = to_device(d_tbase, tb, s)) return r;
    if (int r = to_device(d_counts, cnt, s)) return r;
    if (int r = d_toff.alloc((t + 1) * 4, s)) return r;
    if (int r = d_payload.alloc((n + 1) * 8, s)) return r;
    uint64_t max_t = 0;
    for (uint32_t k = 0; k < n; ++k) max_t = std::max<uint64_t>(max_t, tb[k + 1] - tb[k]);
    // one block row per chunk when the rows of the widest chunk waste at most
    // about as many threads as there are tuples
    const bool g2 = n > 0 && n <= 65535 && ((max_t + 255) / 256) * 256 * n <= 2 * t + 256ull * n;
    pl = SviPlan{d_starts_sel.as<uint64_t>(), d_tbase.as<uint64_t>(), d_counts.as<uint32_t>(), n, t,
                 (uint32_t)max_t, g2 ? 1 : 0};
    if (t) hipLaunchKernelGGL(k_svi_tuple_sizes, svi_grid(pl), dim3(256), 0, s, d_values, pl, d_toff.as<uint32_t>());
    if (n) hipLaunchKernelGGL(k_chunk_scan_u32<256>, dim3(n), dim3(256), 0, s, d_toff.as<uint32_t>(), pl.tuple_base,
                              d_payload.as<uint64_t>());
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint64_t> payload;
    if (int r = to_host(payload, d_payload.p, n, s)) return r;
    chunk_bytes.resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        if (cnt[k] == 0) {
            chunk_bytes[k] = 0;  // StreamVInt::Encode writes nothing for an empty chunk
            continue;
        }
        const uint64_t adjusted = (cnt[k] + 7) / 8 * 8;
        chunk_bytes[k] = 4 + payload[k] + (d_bounds ? adjusted / 2 : 0);
    }
    return PS_OK;
}

int svi_write(const uint32_t* d_values, const uint8_t* d_bounds, const SviPlan& pl, DevBuf& d_toff, DevBuf& d_payload,
              const std::vector<uint64_t>& out_off_sel, uint8_t* d_out, hipStream_t s) {
    DevBuf d_off;
    if (int r = to_device(d_off, out_off_sel, s)) return r;
    if (pl.ntuples)
        hipLaunchKernelGGL(k_svi_write, svi_grid(pl), dim3(256), 0, s, d_values, d_bounds, pl,
                           d_toff.as<uint32_t>(), d_off.as<uint64_t>(), d_payload.as<uint64_t>(), d_out);
    PS_HIP_CHECK(hipGetLastError());
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
}

int read_starts(std::vector<uint64_t>& starts, const uint64_t* d_starts, uint32_t nchunks, hipStream_t s) {
    if (int r = to_host(starts, d_starts, (size_t)nchunks + 1, s)) return r;
    for (uint32_t c = 0; c < nchunks; ++c)
        if (starts[c + 1] < starts[c]) return fail(PS_ERR_INVALID_ARGUMENT, "chunk starts must not decrease");
    return PS_OK;
}

// Decodes the SVI chunks `sel` of a concatenated stream into values.
int svi_decode_chunks(const uint8_t* d_in, const uint64_t* d_in_off, const std::vector<uint32_t>& sel,
                      const std::vector<uint32_t>& counts_all, const uint32_t* d_counts_all,
                      const uint64_t* d_value_starts, uint32_t* d_values, uint8_t* d_bounds, hipStream_t s) {
    const uint32_t n = (uint32_t)sel.size();
    if (n == 0) return PS_OK;
    std::vector<uint64_t> tb(n + 1);
    uint64_t t = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t cnt = counts_all[sel[k]];
        if (cnt > MAX_CHUNK_VALUES) return fail(PS_ERR_INVALID_ARGUMENT, "stream chunk claims %u values", cnt);
        tb[k] = t;
        t += (uint64_t)((cnt + 7) / 8 * 8) / 4;
    }
    tb[n] = t;
    DevBuf d_sel, d_tb, d_toff, d_end, d_deltas, d_err;
    if (int r = to_device(d_sel, sel, s)) return r;
    if (int r = to_device(d_tb, tb, s)) return r;
    if (int r = d_toff.alloc((t + 1) * 4, s)) return r;
    if (int r = d_end.alloc((size_t)n * 4, s)) return r;
    if (int r = d_deltas.alloc((t + 1) * 16, s)) return r;
    const std::vector<uint32_t> zero(1, 0u);
    if (int r = to_device(d_err, zero, s)) return r;
    hipLaunchKernelGGL(k_svi_walk_par<256>, dim3(n), dim3(256), 0, s, d_in, d_in_off, d_tb.as<uint64_t>(),
                       d_sel.as<uint32_t>(), d_bounds ? 1 : 0, d_toff.as<uint32_t>(), d_end.as<uint32_t>(),
                       d_err.as<uint32_t>());
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> bad;
    if (int r = to_host(bad, d_err.p, 1, s)) return r;
    if (bad[0])
        return fail(PS_ERR_INVALID_ARGUMENT, "%u StreamVInt chunks do not end where their tuples end", bad[0]);
    if (t) {
        uint64_t max_t = 0;
        for (uint32_t k = 0; k < n; ++k) max_t = std::max<uint64_t>(max_t, tb[k + 1] - tb[k]);
        const bool g2 = n <= 65535 && ((max_t + 255) / 256) * 256 * n <= 2 * t + 256ull * n;
        const dim3 grid = g2 ? dim3((uint32_t)((max_t + 255) / 256), n) : dim3((uint32_t)blocks_for(t, 256));
        hipLaunchKernelGGL(k_svi_decode_tuples, grid, dim3(256), 0, s, d_in, d_in_off, d_tb.as<uint64_t>(),
                           d_sel.as<uint32_t>(), n, t, g2 ? 1 : 0, d_toff.as<uint32_t>(), d_deltas.as<uint32_t>());
    }
    hipLaunchKernelGGL(k_svi_finish<256>, dim3(n), dim3(256), 0, s, d_deltas.as<uint32_t>(), d_tb.as<uint64_t>(),
                       d_counts_all, d_sel.as<uint32_t>(), d_value_starts, d_values, d_in, d_in_off,
                       d_end.as<uint32_t>(), d_bounds);
    PS_HIP_CHECK(hipGetLastError());
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
}

}  // namespace

int svi_encode_chunks(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
                      uint8_t* d_out, uint64_t out_cap, std::vector<uint64_t>& off, hipStream_t s) {
    const uint32_t nchunks = (uint32_t)(starts.size() - 1);
    std::vector<uint32_t> sel(nchunks);
    for (uint32_t c = 0; c < nchunks; ++c) sel[c] = c;
    DevBuf d_st, d_tb, d_cnt, d_toff, d_pay;
    std::vector<uint64_t> bytes;
    SviPlan pl{};
    if (int r = svi_sizes(d_values, d_bounds, starts, sel, s, d_st, d_tb, d_cnt, d_toff, d_pay, bytes, pl)) return r;
    off.assign(nchunks + 1, 0);
    for (uint32_t c = 0; c < nchunks; ++c) off[c + 1] = off[c] + bytes[c];
    if (off[nchunks] > out_cap)
        return fail(PS_ERR_CAPACITY, "encoded size %llu exceeds output capacity %llu",
                    (unsigned long long)off[nchunks], (unsigned long long)out_cap);
    std::vector<uint64_t> off_sel(off.begin(), off.end() - 1);
    return svi_write(d_values, d_bounds, pl, d_toff, d_pay, off_sel, d_out, s);
}

int fc_encode_chunks(const uint32_t* d_values, const std::vector<uint64_t>& starts, int with_check, uint8_t* d_out,
                     uint64_t out_cap, std::vector<uint64_t>& off, uint64_t* d_offsets, hipStream_t s) {
    const uint32_t nchunks = (uint32_t)(starts.size() - 1);
    DevBuf d_starts, d_first, d_last, d_ok, d_off_own;
    if (int r = to_device(d_starts, starts, s)) return r;
    if (!d_offsets) {
        if (int r = d_off_own.alloc(((size_t)nchunks + 1) * 8, s)) return r;
        d_offsets = d_off_own.as<uint64_t>();
    }
    if (int r = d_first.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (int r = d_last.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (int r = d_ok.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (nchunks)
        hipLaunchKernelGGL(k_fc_inspect<1024>, dim3(nchunks), dim3(1024), 0, s, d_values, d_starts.as<uint64_t>(),
                           d_first.as<uint32_t>(), d_last.as<uint32_t>(), d_ok.as<uint32_t>());
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> first, last, ok;
    if (int r = to_host(first, d_first.p, nchunks, s)) return r;
    if (int r = to_host(last, d_last.p, nchunks, s)) return r;
    if (int r = to_host(ok, d_ok.p, nchunks, s)) return r;
    // Common/FrontierCompression.cpp:92-115: bitmap when count > 128 and the
    // bitmap has fewer bits than 10 per value (minus 8 without the check)
    std::vector<uint32_t> bm, sv;
    std::vector<uint64_t> bm_words;
    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint64_t count = starts[c + 1] - starts[c];
        bool use_bitmap = false;
        if (count > 128) {
            const uint64_t min_bits = count * 10;
            const uint64_t f = first[c], l = last[c];
            const uint64_t bits = (l < f) ? ~0ull : ((l - f + 63) / 64 * 8 + 8) * 8;
            if (with_check) use_bitmap = bits < min_bits && ok[c];
            else use_bitmap = bits < min_bits - 8 && l >= f;
            if (!with_check && use_bitmap && !ok[c])
                return fail(PS_ERR_INVALID_ARGUMENT, "chunk %u is not sorted; use the checked encoder", c);
        }
        if (use_bitmap) {
            bm.push_back(c);
            bm_words.push_back(((uint64_t)last[c] - first[c] + 63) / 64);
        } else {
            sv.push_back(c);
        }
    }
    DevBuf d_st, d_tb, d_cnt, d_toff, d_pay;
    std::vector<uint64_t> sv_bytes;
    SviPlan pl{};
    if (int r = svi_sizes(d_values, nullptr, starts, sv, s, d_st, d_tb, d_cnt, d_toff, d_pay, sv_bytes, pl)) return r;
    std::vector<uint64_t> bytes(nchunks, 0);
    for (size_t k = 0; k < sv.size(); ++k) bytes[sv[k]] = sv_bytes[k];
    for (size_t k = 0; k < bm.size(); ++k) bytes[bm[k]] = 8 + bm_words[k] * 8;
    off.assign(nchunks + 1, 0);
    for (uint32_t c = 0; c < nchunks; ++c) off[c + 1] = off[c] + bytes[c];
    if (off[nchunks] > out_cap)
        return fail(PS_ERR_CAPACITY, "encoded size %llu exceeds output capacity %llu",
                    (unsigned long long)off[nchunks], (unsigned long long)out_cap);
    PS_HIP_CHECK(hipMemcpyAsync(d_offsets, off.data(), off.size() * 8, hipMemcpyHostToDevice, s));
    std::vector<uint64_t> sv_off(sv.size());
    for (size_t k = 0; k < sv.size(); ++k) sv_off[k] = off[sv[k]];
    if (int r = svi_write(d_values, nullptr, pl, d_toff, d_pay, sv_off, d_out, s)) return r;
    if (!bm.empty()) {
        std::vector<uint64_t> wb(bm.size() + 1, 0);
        for (size_t k = 0; k < bm.size(); ++k) wb[k + 1] = wb[k] + bm_words[k];
        DevBuf d_sel, d_wb, d_words;
        if (int r = to_device(d_sel, bm, s)) return r;
        if (int r = to_device(d_wb, wb, s)) return r;
        if (int r = d_words.alloc(wb.back() * 8 + 8, s)) return r;
        if (int r = launch_fill(d_words.p, 0, wb.back() * 8 + 8, s)) return r;
        hipLaunchKernelGGL(k_fc_bitmap_set, dim3((uint32_t)bm.size()), dim3(256), 0, s, d_values,
                           d_starts.as<uint64_t>(), d_sel.as<uint32_t>(), (uint32_t)bm.size(), d_wb.as<uint64_t>(),
                           d_first.as<uint32_t>(), d_words.as<unsigned long long>());
        hipLaunchKernelGGL(k_fc_bitmap_write, dim3((uint32_t)bm.size()), dim3(256), 0, s, d_sel.as<uint32_t>(),
                           (uint32_t)bm.size(), d_wb.as<uint64_t>(), d_first.as<uint32_t>(),
                           d_words.as<unsigned long long>(), d_offsets, d_out);
        PS_HIP_CHECK(hipGetLastError());
        PS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return PS_OK;
}

}  // namespace ps

using namespace ps;

extern "C" {

int ps_svint_encode(const uint32_t* d_values, const uint8_t* d_bounds, const uint64_t* d_starts, uint32_t nchunks,
                    uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_starts || !d_offsets || (nchunks && (!d_values || !d_out)))
        return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipStream_t s = stream_of(st);
    std::vector<uint64_t> starts, off;
    if (int r = read_starts(starts, d_starts, nchunks, s)) return r;
    if (int r = svi_encode_chunks(d_values, d_bounds, starts, d_out, out_cap, off, s)) return r;
    PS_HIP_CHECK(hipMemcpyAsync(d_offsets, off.data(), off.size() * 8, hipMemcpyHostToDevice, s));
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
    PS_API_END
}

int ps_svint_decode(const uint8_t* d_in, const uint64_t* d_in_offsets, uint32_t nchunks, int32_t with_bounds,
                    uint32_t* d_values, uint8_t* d_bounds, uint64_t values_cap, uint64_t* d_value_starts,
                    ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_in_offsets || !d_value_starts || (nchunks && (!d_in || !d_values)) || (with_bounds && !d_bounds))
        return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipStream_t s = stream_of(st);
    DevBuf d_counts;
    if (int r = d_counts.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (nchunks)
        hipLaunchKernelGGL(k_read_counts, dim3(blocks_for(nchunks, 256)), dim3(256), 0, s, d_in, d_in_offsets, nchunks,
                           d_counts.as<uint32_t>(), (uint32_t*)nullptr);
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> counts;
    if (int r = to_host(counts, d_counts.p, nchunks, s)) return r;
    std::vector<uint64_t> vs(nchunks + 1, 0);
    for (uint32_t c = 0; c < nchunks; ++c) vs[c + 1] = vs[c] + counts[c];
    if (vs[nchunks] > values_cap)
        return fail(PS_ERR_CAPACITY, "%llu decoded values exceed capacity %llu", (unsigned long long)vs[nchunks],
                    (unsigned long long)values_cap);
    PS_HIP_CHECK(hipMemcpyAsync(d_value_starts, vs.data(), vs.size() * 8, hipMemcpyHostToDevice, s));
    std::vector<uint32_t> sel;
    for (uint32_t c = 0; c < nchunks; ++c)
        if (counts[c]) sel.push_back(c);
    return svi_decode_chunks(d_in, d_in_offsets, sel, counts, d_counts.as<uint32_t>(), d_value_starts, d_values,
                             with_bounds ? d_bounds : nullptr, s);
    PS_API_END
}

int ps_opbits_encode(int32_t bits, const uint8_t* d_ops, uint64_t count, uint8_t* d_out, ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bits != 2 && bits != 4 && bits != 8) return fail(PS_ERR_INVALID_ARGUMENT, "bits must be 2, 4 or 8");
    if (count % 8) return fail(PS_ERR_INVALID_ARGUMENT, "count must be a multiple of 8");
    if (count == 0) return PS_OK;
    if (!d_ops || !d_out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    const uint64_t groups = count / 8;
    hipLaunchKernelGGL(k_opbits_encode, dim3(blocks_for(groups, 256)), dim3(256), 0, stream_of(st), bits, d_ops,
                       groups, d_out);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

int ps_opbits_decode(int32_t bits, const uint8_t* d_in, uint64_t count, uint8_t* d_ops, ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (bits != 2 && bits != 4 && bits != 8) return fail(PS_ERR_INVALID_ARGUMENT, "bits must be 2, 4 or 8");
    if (count % 8) return fail(PS_ERR_INVALID_ARGUMENT, "count must be a multiple of 8");
    if (count == 0) return PS_OK;
    if (!d_ops || !d_in) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    const uint64_t groups = count / 8;
    hipLaunchKernelGGL(k_opbits_decode, dim3(blocks_for(groups, 256)), dim3(256), 0, stream_of(st), bits, d_in,
                       groups, d_ops);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

int ps_fc_encode(const uint32_t* d_values, const uint64_t* d_starts, uint32_t nchunks, int32_t with_check,
                 uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_starts || !d_offsets || (nchunks && (!d_values || !d_out)))
        return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipStream_t s = stream_of(st);
    std::vector<uint64_t> starts, off;
    if (int r = read_starts(starts, d_starts, nchunks, s)) return r;
    if (int r = fc_encode_chunks(d_values, starts, with_check, d_out, out_cap, off, d_offsets, s)) return r;
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
    PS_API_END
}

int ps_fc_decode(const uint8_t* d_in, const uint64_t* d_in_offsets, uint32_t nchunks, uint32_t* d_values,
                 uint64_t values_cap, uint64_t* d_value_starts, ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!d_in_offsets || !d_value_starts || (nchunks && (!d_in || !d_values)))
        return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipStream_t s = stream_of(st);
    DevBuf d_counts, d_isbm;
    if (int r = d_counts.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (int r = d_isbm.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (nchunks)
        hipLaunchKernelGGL(k_read_counts, dim3(blocks_for(nchunks, 256)), dim3(256), 0, s, d_in, d_in_offsets, nchunks,
                           d_counts.as<uint32_t>(), d_isbm.as<uint32_t>());
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> isbm;
    if (int r = to_host(isbm, d_isbm.p, nchunks, s)) return r;
    std::vector<uint32_t> bm, sv;
    for (uint32_t c = 0; c < nchunks; ++c) (isbm[c] ? bm : sv).push_back(c);
    DevBuf d_bm, d_words;
    if (!bm.empty()) {
        // bitmap chunks: words in the header; the count is 1 + popcount(bitmap)
        std::vector<uint32_t> words_all;
        if (int r = to_host(words_all, d_counts.p, nchunks, s)) return r;
        if (int r = to_device(d_words, words_all, s)) return r;
        if (int r = to_device(d_bm, bm, s)) return r;
        hipLaunchKernelGGL(k_fc_bitmap_count, dim3((uint32_t)bm.size()), dim3(256), 0, s, d_in, d_in_offsets,
                           d_bm.as<uint32_t>(), (uint32_t)bm.size(), d_words.as<uint32_t>(), d_counts.as<uint32_t>());
        PS_HIP_CHECK(hipGetLastError());
    }
    std::vector<uint32_t> counts;
    if (int r = to_host(counts, d_counts.p, nchunks, s)) return r;
    std::vector<uint64_t> vs(nchunks + 1, 0);
    for (uint32_t c = 0; c < nchunks; ++c) vs[c + 1] = vs[c] + counts[c];
    if (vs[nchunks] > values_cap)
        return fail(PS_ERR_CAPACITY, "%llu decoded values exceed capacity %llu", (unsigned long long)vs[nchunks],
                    (unsigned long long)values_cap);
    PS_HIP_CHECK(hipMemcpyAsync(d_value_starts, vs.data(), vs.size() * 8, hipMemcpyHostToDevice, s));
    if (!bm.empty()) {
        hipLaunchKernelGGL(k_fc_bitmap_decode, dim3((uint32_t)bm.size()), dim3(64), 0, s, d_in, d_in_offsets,
                           d_bm.as<uint32_t>(), (uint32_t)bm.size(), d_words.as<uint32_t>(), d_value_starts, d_values);
        PS_HIP_CHECK(hipGetLastError());
    }
    std::vector<uint32_t> sv_nonempty;
    for (uint32_t c : sv)
        if (counts[c]) sv_nonempty.push_back(c);
    return svi_decode_chunks(d_in, d_in_offsets, sv_nonempty, counts, d_counts.as<uint32_t>(), d_value_starts, d_values,
                             nullptr, s);
    PS_API_END
}

int ps_bucket_by_segment(const uint64_t* d_children, const uint8_t* d_ops, uint64_t count, int32_t segment_bits,
                         uint32_t nseg, int32_t nops, uint32_t* d_sorted_idx, uint64_t* d_bucket_starts,
                         ps_stream st) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (segment_bits < 1 || segment_bits > 32) return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits must be 1..32");
    if (nops < 1 || nops > 16) return fail(PS_ERR_INVALID_ARGUMENT, "nops must be 1..16");
    if (nseg == 0 || (uint64_t)nseg * nops > (1ull << 24)) return fail(PS_ERR_INVALID_ARGUMENT, "bad segment count");
    if (count > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "at most 2^31-1 children per call");
    if (!d_bucket_starts || (count && (!d_children || !d_ops || !d_sorted_idx)))
        return fail(PS_ERR_INVALID_ARGUMENT, "NULL device pointer");
    hipStream_t s = stream_of(st);
    const uint64_t nb = (uint64_t)nseg * nops;
    std::vector<uint64_t> starts(nb + 1, 0);
    if (count) {
        // (op * nseg + segment) << 32 | in-segment index, sorted by index with
        // 8-bit stable counting passes, then stably grouped by bucket: buckets
        // in (op, segment) order, each ascending
        DevBuf d_kv, d_tmp;
        if (int r = d_kv.alloc(count * 8, s)) return r;
        if (int r = d_tmp.alloc(count * 8, s)) return r;
        hipLaunchKernelGGL(k_bucket_keys, dim3(blocks_for(count, 256)), dim3(256), 0, s, d_children, d_ops, count,
                           segment_bits, nseg, (uint32_t)nops, d_kv.as<unsigned long long>());
        PS_HIP_CHECK(hipGetLastError());
        if (int r = stable_sort_by_value(d_kv.as<unsigned long long>(), d_tmp.as<unsigned long long>(), count,
                                         segment_bits, s))
            return r;
        std::vector<uint32_t> keys;
        std::vector<uint64_t> counts;
        uint64_t dropped = 0;
        if (int r = stable_group_by_key(d_kv.as<unsigned long long>(), count, (uint32_t)nb, d_sorted_idx, keys, counts,
                                        &dropped, s))
            return r;
        if (dropped) return fail(PS_ERR_INDEX_RANGE, "children outside [0, nseg) segments or ops >= nops");
        for (size_t k = 0; k < keys.size(); ++k) starts[keys[k] + 1] = counts[k];
        for (uint64_t b = 0; b < nb; ++b) starts[b + 1] += starts[b];
    }
    PS_HIP_CHECK(hipMemcpyAsync(d_bucket_starts, starts.data(), starts.size() * 8, hipMemcpyHostToDevice, s));
    PS_HIP_CHECK(hipStreamSynchronize(s));
    return PS_OK;
    PS_API_END
}

}  // extern "C"
