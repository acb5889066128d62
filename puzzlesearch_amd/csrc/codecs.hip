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
// per chunk for byte offsets, then the writes.  Decoding is one pass per chunk
// (k_svi_decode_chunk): LDS-staged tiles, a parallel walk of the tuple
// descriptors composed by a block scan, the deltas undone by a second scan.
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

// skip (optional): chunks another encoder writes (bitmap chunks); their sizes,
// and so their payload sums, are left undefined and unused
__global__ void k_svi_tuple_sizes(const uint32_t* v, SviPlan pl, uint32_t* sizes, const uint32_t* skip) {
    uint64_t t;
    uint32_t c;
    if (!svi_tuple(pl, t, c)) return;
    if (skip && skip[c]) return;
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

// skip (optional): chunks written by another encoder (bitmap chunks)
// out_cap != ~0: chunk_out holds nchunks + 1 offsets, and a chunk ending past
// out_cap is not written (counted in *err by its first tuple's thread)
__global__ void k_svi_write(const uint32_t* v, const uint8_t* bounds, SviPlan pl, const uint32_t* tuple_off,
                            const uint64_t* chunk_out, const uint64_t* chunk_payload, uint8_t* out,
                            const uint32_t* skip, uint64_t out_cap, uint32_t* err) {
    uint64_t t;
    uint32_t c;
    if (!svi_tuple(pl, t, c)) return;
    if (skip && skip[c]) return;
    if (out_cap != ~0ull && chunk_out[c + 1] > out_cap) {
        if (err && t == pl.tuple_base[c]) atomicAdd(err, 1u);
        return;
    }
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

// One pass per chunk: a block decodes one StreamVInt chunk, its payload
// [4, E) (E = chunk bytes - bound bytes) in tiles of SVD_TILE bytes staged in
// LDS (with SVD_LOOK bytes past the tile, where a tuple starting in the tile
// may end).  A tuple is a descriptor byte and four values of 1..4 bytes
// (5..17 bytes), so where the tuples of a stretch of bytes start depends on
// every earlier descriptor:
//  1 thread t owns the tile's bytes [ts + t*B, ts + (t+1)*B) and walks them from
//    all 17 offsets at which the first tuple starting in them can begin,
//    recording for each where the walk leaves its bytes (0..16 past their end)
//    and how many tuples start in them: a map entry offset -> (exit, tuples);
//  2 a block scan composes the maps (Hillis-Steele over the threads, 17
//    entries each), so every thread learns, from the tile's entry offset, the
//    offset of its first tuple and the tuple's index in the tile;
//  3 every thread decodes its tuples' deltas into LDS, a block scan of the
//    per-thread sums undoes the delta coding (the running value carries from
//    tile to tile), and the tile's values are stored with coalesced writes;
//  4 the last tile must end exactly at E, after exactly T = ceil(count/8)*2
//    tuples; a chunk that does not (or whose payload length no T tuples can
//    fill) is refused and counted in *err, its values undefined.
constexpr int SVD_TPB = 512;
constexpr uint32_t SVD_B = 32;                        // payload bytes per thread and tile
constexpr uint32_t SVD_TILE = SVD_B * SVD_TPB;        // 16 KB
constexpr uint32_t SVD_LOOK = 16;                     // bytes past the tile a tuple can reach
constexpr uint32_t SVD_NE = 17;                       // entry offsets 0..16
constexpr uint32_t SVD_MAXT = (SVD_TILE + 4) / 5;     // tuples starting in one tile, at most
// staged bytes are laid out with a pad word after every 8 words, so the
// 32-byte stretches of consecutive threads start in different LDS banks
constexpr uint32_t SVD_STAGE_WORDS = (SVD_TILE + SVD_LOOK + 4) / 4;
constexpr uint32_t SVD_PAD_WORDS = (SVD_STAGE_WORDS + 7) / 8 * 9 + 1;
constexpr size_t SVD_LDS = (size_t)SVD_PAD_WORDS * 4 + (size_t)2 * SVD_TPB * SVD_NE * 4 + (size_t)SVD_MAXT * 16;

__device__ __forceinline__ uint32_t svd_pw(uint32_t w) { return (w >> 3) * 9u + (w & 7u); }
__device__ __forceinline__ uint32_t svd_pb(uint32_t b) { return (b >> 5) * 36u + (b & 31u); }

__device__ __forceinline__ uint32_t svd_len(uint32_t d) {
    // 5 + the four 2-bit length fields
    return 5u + (uint32_t)__popc(d & 0x55u) + 2u * (uint32_t)__popc(d & 0xAAu);
}

// exclusive block sum of v over SVD_TPB threads; *total receives the sum
__device__ __forceinline__ uint32_t svd_block_excl(uint32_t v, uint32_t* s_w, uint32_t* total) {
    constexpr int NW = SVD_TPB / 64;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t incl = wave_incl_scan_u32(v, lane);
    __syncthreads();
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t x = s_w[w];
        before += (w < wave) ? x : 0u;
        tot += x;
    }
    *total = tot;
    return before + incl - v;
}

__global__ __launch_bounds__(SVD_TPB) void k_svi_decode_chunk(const uint8_t* in, const uint64_t* in_off,
                                                               const uint32_t* select, const uint64_t* value_starts,
                                                               int with_bounds, uint32_t* out, uint8_t* bounds_out,
                                                               uint32_t* err, uint64_t values_cap) {
    extern __shared__ uint32_t svd_smem[];
    uint8_t* buf = reinterpret_cast<uint8_t*>(svd_smem);  // staged bytes, from an aligned address
    uint32_t* map0 = svd_smem + SVD_PAD_WORDS;
    uint32_t* map1 = map0 + SVD_TPB * SVD_NE;
    uint32_t* vals = map1 + SVD_TPB * SVD_NE;
    __shared__ uint32_t s_w[SVD_TPB / 64];
    __shared__ uint32_t s_bad;
    // without `select` every chunk gets a block: empty chunks are skipped, and
    // a chunk past values_cap (err[1]) or over MAX_CHUNK_VALUES (err[2]) is refused
    const uint32_t c = select ? select[blockIdx.x] : blockIdx.x;
    if (!select) {
        const uint64_t v0 = value_starts[c], v1 = value_starts[c + 1];
        if (v1 == v0) return;
        if (v1 - v0 > MAX_CHUNK_VALUES || v1 > values_cap) {
            if (threadIdx.x == 0) atomicAdd(err + (v1 - v0 > MAX_CHUNK_VALUES ? 2 : 1), 1u);
            return;
        }
    }
    const uint8_t* p = in + in_off[c];
    const uint64_t clen = in_off[c + 1] - in_off[c];
    const int tid = threadIdx.x;
    if (tid == 0) s_bad = 0;
    __syncthreads();
    if (clen < 4) {
        if (tid == 0) atomicAdd(err, 1u);
        return;
    }
    const uint32_t count = rd32(p);
    const uint64_t T = (uint64_t)((count + 7u) / 8u) * 2u;  // tuples: count padded to a multiple of 8, 4 per tuple
    const uint64_t bnd = with_bounds ? T * 2 : 0;
    const uint64_t E = clen >= 4 + bnd ? clen - bnd : 4;
    if (clen < 4 + bnd || E - 4 < 5 * T || E - 4 > 17 * T) {
        if (tid == 0) atomicAdd(err, 1u);
        return;
    }
    uint32_t* o = out + value_starts[c];
    uint32_t entry = 0;   // offset of the tile's first tuple past the tile start
    uint64_t tbase = 0;   // tuples before the tile
    uint32_t carry = 0;   // running value before the tile
    // the staged words of a tile: [ts, min(te + LOOK, clen)) from the aligned
    // address below ts, bytes outside the chunk as 0; the next tile's words
    // are loaded into registers while the current tile is decoded
    constexpr uint32_t NWORDS = SVD_STAGE_WORDS;
    constexpr uint32_t WPT = (NWORDS + SVD_TPB - 1) / SVD_TPB;
    auto load_tile = [&](uint64_t ts, uint32_t (&r)[WPT]) {
        const uint64_t te = ts + SVD_TILE < E ? ts + SVD_TILE : E;
        const uint64_t se = te + SVD_LOOK < clen ? te + SVD_LOOK : clen;
        const uintptr_t lo = reinterpret_cast<uintptr_t>(p + ts), hi = reinterpret_cast<uintptr_t>(p + se);
        const uintptr_t a0 = lo & ~(uintptr_t)3;
#pragma unroll
        for (uint32_t k = 0; k < WPT; ++k) {
            const uint32_t w = tid + k * SVD_TPB;
            const uintptr_t a = a0 + (uintptr_t)w * 4;
            uint32_t x = 0;
            if (w < NWORDS && a + 4 <= hi && a >= lo) {
                x = *reinterpret_cast<const uint32_t*>(a);
            } else if (w < NWORDS) {
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (a + b >= lo && a + b < hi) x |= (uint32_t)*reinterpret_cast<const uint8_t*>(a + b) << (8 * b);
            }
            r[k] = x;
        }
    };
    uint32_t staged[WPT];
    // Every delta after the first of one byte (dense frontiers; the first is
    // the chunk's first value): when the payload is the first tuple and then
    // exactly 5 bytes per tuple, and those descriptors are all 0, tuple t >= 1
    // starts at 4 + L0 + 5(t-1), so no walk is needed.  Each step stages
    // 512 x 6 tuples (15360 bytes) and every thread decodes six; a nonzero
    // descriptor sends the chunk to the walk below, which rewrites its values.
    const uint32_t L0 = svd_len(p[4]);
    if (4 + (uint64_t)L0 <= E && E - 4 - L0 == 5 * (T - 1)) {
        constexpr uint32_t K = 6;
        constexpr uint64_t STEP = (uint64_t)SVD_TPB * K;
        const uint64_t T1 = T - 1;
        uint32_t carry5 = 0;
        {  // the first tuple, read directly
            const uint32_t d = p[4];
            uint32_t k = 5, run = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t len = 1u + ((d >> (2 * j)) & 3u);
                uint32_t x = 0;
                for (uint32_t i = 0; i < len; ++i) x |= (uint32_t)p[k + i] << (8 * i);
                k += len;
                run += x;
                if (tid == 0 && (uint32_t)j < count) o[j] = run;
            }
            carry5 = run;
        }
        bool bad = false;
        load_tile(4 + L0, staged);
        for (uint64_t t0 = 0; t0 < T1; t0 += STEP) {
            const uint64_t ts = 4 + L0 + 5 * t0;
            const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(p + ts) & 3u);
#pragma unroll
            for (uint32_t k = 0; k < WPT; ++k) {
                const uint32_t w = tid + k * SVD_TPB;
                if (w < NWORDS) svd_smem[svd_pw(w)] = staged[k];
            }
            __syncthreads();
            if (t0 + STEP < T1) load_tile(4 + L0 + 5 * (t0 + STEP), staged);
            // my six tuples: bytes [30 tid, 30 tid + 30) of the staged step
            const uint32_t b0 = sh + 30u * tid;
            uint32_t wv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) wv[k] = svd_smem[svd_pw((b0 >> 2) + k)];
            const uint32_t r8 = (b0 & 3u) * 8u;
            uint32_t v[K * 4];
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t j = 0; j < K; ++j) {
                // bytes 5j .. 5j+4 of my stretch: descriptor, then four one-byte deltas
                const uint32_t o0 = r8 + 40u * j;                    // bit offset of the descriptor
                const uint64_t lo = (uint64_t)wv[o0 >> 5] | ((uint64_t)wv[(o0 >> 5) + 1] << 32);
                const uint64_t x = lo >> (o0 & 31u);                // descriptor in bits 0..7
                const bool live = t0 + (uint64_t)tid * K + j < T1;
                bad |= live && (x & 255u) != 0u;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const uint32_t d = live ? (uint32_t)(x >> (8 * (e + 1))) & 255u : 0u;
                    sum += d;
                    v[j * 4 + e] = sum;
                }
            }
            uint32_t tot;
            const uint32_t base = carry5 + svd_block_excl(sum, s_w, &tot);
            const uint64_t j0 = (1 + t0 + (uint64_t)tid * K) * 4;
#pragma unroll
            for (uint32_t e = 0; e < K * 4; ++e)
                if (j0 + e < count) o[j0 + e] = base + v[e];
            carry5 += tot;
            __syncthreads();  // the staged bytes are rewritten by the next step
        }
        if (bad) s_bad = 1;
        __syncthreads();
        if (!s_bad) {
            if (bounds_out) {
                const uint8_t* q = p + E;
                uint8_t* bo = bounds_out + value_starts[c];
                for (uint64_t g = tid; g * 8 < count; g += SVD_TPB) {
                    const uint64_t stored = rd32(q + g * 4);
                    const uint64_t all = (stored & 0x0F0F0F0Full) | ((stored & 0xF0F0F0F0ull) << 28);
                    for (int e = 0; e < 8; ++e) {
                        const uint64_t idx = g * 8 + e;
                        if (idx < count) bo[idx] = (uint8_t)(all >> (8 * e));
                    }
                }
            }
            return;
        }
        __syncthreads();
        if (tid == 0) s_bad = 0;
        __syncthreads();
    }
    load_tile(4, staged);
    for (uint64_t ts = 4; ts < E; ts += SVD_TILE) {
        const uint64_t te = ts + SVD_TILE < E ? ts + SVD_TILE : E;
        const uint32_t sh = (uint32_t)(reinterpret_cast<uintptr_t>(p + ts) & 3u);
#pragma unroll
        for (uint32_t k = 0; k < WPT; ++k) {
            const uint32_t w = tid + k * SVD_TPB;
            if (w < NWORDS) svd_smem[svd_pw(w)] = staged[k];
        }
        __syncthreads();
        if (te < E) load_tile(te, staged);
        // 1 the walk map of this thread's bytes [rs, re), positions relative to ts
        const uint32_t tl = (uint32_t)(te - ts);
        const uint32_t rs = tid * SVD_B < tl ? tid * SVD_B : tl;
        const uint32_t re = rs + SVD_B < tl ? rs + SVD_B : tl;
        // A tile whose tuples from the entry on are all of 5 bytes (every
        // descriptor on the chain entry, entry+5, .. is 0) needs no walk: the
        // block checks the bytes of the entry's residue class mod 5 at or
        // after the entry.
        __shared__ uint32_t s_nz;
        if (tid == 0) s_nz = 0u;
        __syncthreads();
        {
            const uint32_t r5 = entry % 5u;
            bool nz = false;
            for (uint32_t i = rs + (r5 + 5u - rs % 5u) % 5u; i < re; i += 5u)
                nz |= i >= entry && buf[svd_pb(sh + i)] != 0u;
            if (__ballot(nz) != 0ull && (tid % 64) == 0) s_nz = 1u;
        }
        __syncthreads();
        const bool uniform = entry < tl && s_nz == 0u;
        uint32_t my_entry, my_t, tile_tuples, next_entry;
        if (uniform) {
            // chain positions entry + 5k: the first at or after rs, and how many come before it
            const uint32_t k0 = rs > entry ? (rs - entry + 4u) / 5u : 0u;
            my_entry = entry + 5u * k0 - rs;
            my_t = k0;
            tile_tuples = (tl - entry + 4u) / 5u;
            next_entry = entry + 5u * tile_tuples - tl;
        } else {
            {
                // from the last byte back: f(i) = where the walk from a tuple
                // starting at byte i leaves [rs, re) and how many tuples start
                // on the way, f(i) = f(i + len(i)) plus one tuple; kept in the
                // value buffer, free until the tuples are decoded
                uint16_t* fd = reinterpret_cast<uint16_t*>(vals);  // [SVD_B][SVD_TPB]
                for (int i = (int)re - 1; i >= (int)rs; --i) {
                    const uint32_t j = (uint32_t)i + svd_len(buf[svd_pb(sh + (uint32_t)i)]);
                    const uint32_t f = j >= re ? ((1u << 8) | (j - re))
                                               : (uint32_t)fd[(j - rs) * SVD_TPB + tid] + (1u << 8);
                    fd[((uint32_t)i - rs) * SVD_TPB + tid] = (uint16_t)f;
                }
#pragma unroll
                for (uint32_t e = 0; e < SVD_NE; ++e) {
                    const uint32_t i = rs + e;
                    map0[tid * SVD_NE + e] = i >= re ? i - re : (uint32_t)fd[e * SVD_TPB + tid];
                }
            }
            __syncthreads();
            // 2 inclusive composition scan: map[t] = the walk of threads 0..t
            uint32_t* A = map0;
            uint32_t* Bm = map1;
            for (int d = 1; d < SVD_TPB; d <<= 1) {
#pragma unroll
                for (uint32_t e = 0; e < SVD_NE; ++e) {
                    uint32_t r = A[tid * SVD_NE + e];
                    if (tid >= d) {
                        const uint32_t f = A[(tid - d) * SVD_NE + e];
                        const uint32_t g = A[tid * SVD_NE + (f & 255u)];
                        r = ((f >> 8) + (g >> 8)) << 8 | (g & 255u);
                    }
                    Bm[tid * SVD_NE + e] = r;
                }
                __syncthreads();
                uint32_t* t = A;
                A = Bm;
                Bm = t;
            }
            const uint32_t before = tid ? A[(tid - 1) * SVD_NE + entry] : entry;  // (tuples << 8) | my entry
            my_entry = tid ? (before & 255u) : entry;
            my_t = tid ? (before >> 8) : 0u;
            const uint32_t last = A[(SVD_TPB - 1) * SVD_NE + entry];
            tile_tuples = last >> 8;
            next_entry = last & 255u;
        }
        // 3 deltas of this thread's tuples into LDS, then the values
        uint32_t sum = 0;
        uint32_t nt = 0;
        for (uint32_t pos = rs + my_entry; pos < re; ++nt) {
            const uint32_t q = my_t + nt;
            // bytes pos .. pos+16 of the tuple from five aligned LDS words
            const uint32_t qb = sh + pos, r8 = (qb & 3u) * 8u;
            const uint32_t wq = qb >> 2;
            const uint64_t w01 = (uint64_t)svd_smem[svd_pw(wq)] | ((uint64_t)svd_smem[svd_pw(wq + 1)] << 32);
            const uint64_t w23 = (uint64_t)svd_smem[svd_pw(wq + 2)] | ((uint64_t)svd_smem[svd_pw(wq + 3)] << 32);
            const uint32_t w4 = svd_smem[svd_pw(wq + 4)];
            const uint64_t b0 = r8 ? (w01 >> r8) | (w23 << (64u - r8)) : w01;                  // bytes 0..7
            const uint64_t b1 = r8 ? (w23 >> r8) | ((uint64_t)w4 << (64u - r8)) : w23;         // bytes 8..15
            const uint32_t b16 = (w4 >> r8) & 255u;                                            // byte 16
            const uint32_t d = (uint32_t)(b0 & 255u);
            const uint64_t xl = (b0 >> 8) | (b1 << 56);             // value bytes 0..7 (tuple bytes 1..8)
            const uint64_t xh = (b1 >> 8) | ((uint64_t)b16 << 56);  // value bytes 8..15
            uint32_t o8 = 0;  // value byte offset
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t len = 1u + ((d >> (2 * j)) & 3u);
                const uint32_t sb = o8 * 8u;
                const uint64_t x = o8 < 8u ? ((xl >> sb) | (sb ? xh << (64u - sb) : 0ull)) : (xh >> (sb - 64u));
                const uint32_t v = (uint32_t)x & (len == 4u ? 0xFFFFFFFFu : ((1u << (8u * len)) - 1u));
                o8 += len;
                if (q < SVD_MAXT) vals[q * 4 + j] = v;
                sum += v;
            }
            pos += 1u + o8;
        }
        uint32_t tile_sum;
        uint32_t run = carry + svd_block_excl(sum, s_w, &tile_sum);
        for (uint32_t i = 0; i < nt && my_t + i < SVD_MAXT; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                run += vals[(my_t + i) * 4 + j];
                vals[(my_t + i) * 4 + j] = run;
            }
        __syncthreads();
        const uint64_t v0 = tbase * 4;  // value index of the tile's first value
        const uint32_t nv = (tile_tuples < SVD_MAXT ? tile_tuples : SVD_MAXT) * 4;
        for (uint32_t i = tid; i < nv; i += SVD_TPB)
            if (v0 + i < count) o[v0 + i] = vals[i];
        if (tile_tuples > SVD_MAXT && tid == 0) s_bad = 1;
        entry = next_entry;
        tbase += tile_tuples;
        carry += tile_sum;
        __syncthreads();  // buf, maps and vals are rewritten by the next tile
    }
    // 4 the tuples end exactly at E, T of them
    if (tid == 0 && (entry != 0 || tbase != T || s_bad)) atomicAdd(err, 1u);
    if (bounds_out) {
        const uint8_t* q = p + E;
        uint8_t* bo = bounds_out + value_starts[c];
        for (uint64_t g = tid; g * 8 < count; g += SVD_TPB) {
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

// ---------------------------------------------------------------- stream-ordered chunk plans

// chunk c's StreamVInt bytes: count, payload and (with bounds) the 4-bit
// bounds of the count rounded up to 8; nothing for an empty chunk
__global__ void k_svi_chunk_bytes(SviPlan pl, const uint64_t* payload, int with_bounds, uint64_t* bytes) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= pl.nchunks) return;
    const uint64_t cnt = pl.counts[c];
    bytes[c] = cnt ? 4 + payload[c] + (with_bounds ? (cnt + 7) / 8 * 8 / 2 : 0) : 0;
}

// FrontierCompression's choice per chunk (Common/FrontierCompression.cpp:92-115):
// a bitmap when count > 128 and the bitmap has fewer bits than 10 per value
// (minus 8 without the check); an unchecked bitmap chunk that is not sorted
// counts in *err.
__global__ void k_fc_kind(SviPlan pl, const uint32_t* first, const uint32_t* last, const uint32_t* ok, int with_check,
                          uint32_t* is_bitmap, uint32_t* words, uint32_t* err) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= pl.nchunks) return;
    const uint64_t count = pl.counts[c];
    bool use_bitmap = false;
    uint64_t nw = 0;
    if (count > 128) {
        const uint64_t min_bits = count * 10;
        const uint64_t f = first[c], l = last[c];
        const uint64_t bits = (l < f) ? ~0ull : ((l - f + 63) / 64 * 8 + 8) * 8;
        if (with_check) use_bitmap = bits < min_bits && ok[c];
        else use_bitmap = bits < min_bits - 8 && l >= f;
        if (!with_check && use_bitmap && !ok[c] && err) atomicAdd(err, 1u);
        if (use_bitmap) nw = (l - f + 63) / 64;
    }
    is_bitmap[c] = use_bitmap ? 1u : 0u;
    words[c] = (uint32_t)nw;
}

// chunk bytes: a bitmap chunk's header and words, a StreamVInt chunk's count
// and payload, nothing for an empty chunk
__global__ void k_fc_bytes(SviPlan pl, const uint64_t* payload, const uint32_t* is_bitmap, const uint32_t* words,
                           uint64_t* bytes) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= pl.nchunks) return;
    const uint64_t count = pl.counts[c];
    bytes[c] = is_bitmap[c] ? 8 + (uint64_t)words[c] * 8 : (count ? 4 + payload[c] : 0);
}

// One block per chunk, bitmap chunks of at most FC_LDS_WORDS words only: the
// bitmap of the values after the first, built in LDS and written after the
// chunk's 8-byte header.
constexpr uint32_t FC_LDS_WORDS = 2048;  // wider bitmaps go through global scratch (k_fc_bitmap_wide_*)

__global__ __launch_bounds__(256) void k_fc_bitmap_chunk(const uint32_t* v, SviPlan pl, const uint32_t* is_bitmap,
                                                         const uint32_t* words, const uint32_t* first,
                                                         const uint64_t* out_off, uint8_t* out, uint64_t out_cap,
                                                         uint32_t* err) {
    const uint32_t c = blockIdx.x;
    if (!is_bitmap[c] || words[c] > FC_LDS_WORDS) return;
    if (out_off[c + 1] > out_cap) {  // not written; counted
        if (threadIdx.x == 0 && err) atomicAdd(err, 1u);
        return;
    }
    __shared__ unsigned long long bits[FC_LDS_WORDS];
    const uint32_t nw = words[c];
    const uint32_t f = first[c];
    const uint64_t b = pl.starts[c], e = b + pl.counts[c];
    for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) bits[i] = 0ull;
    __syncthreads();
    for (uint64_t i = b + 1 + threadIdx.x; i < e; i += blockDim.x) {
        const uint32_t val = v[i] - f - 1;  // in [0, 64 nw): every value lies in (first, last]
        atomicOr(&bits[val >> 6], 1ull << (val & 63));
    }
    __syncthreads();
    uint8_t* o = out + out_off[c];
    if (threadIdx.x == 0) {
        const uint32_t h = nw | CODEC_MAP;
        for (int j = 0; j < 4; ++j) {
            o[j] = (uint8_t)(h >> (8 * j));
            o[4 + j] = (uint8_t)(f >> (8 * j));
        }
    }
    const uint8_t* src = reinterpret_cast<const uint8_t*>(bits);
    for (uint32_t i = threadIdx.x; i < nw * 8; i += blockDim.x) o[8 + i] = src[i];
}

// Bitmap chunks too wide for LDS (more than FC_LDS_WORDS words): the bits are
// ORed into a zeroed global scratch at the chunk's word base, then copied
// after the chunk's header.  One block per chunk; other chunks return.
__global__ __launch_bounds__(256) void k_fc_bitmap_wide_set(const uint32_t* v, SviPlan pl, const uint32_t* is_bitmap,
                                                            const uint32_t* words, const uint32_t* first,
                                                            const unsigned long long* word_base,
                                                            unsigned long long* scratch) {
    const uint32_t c = blockIdx.x;
    if (!is_bitmap[c] || words[c] <= FC_LDS_WORDS) return;
    const uint32_t f = first[c];
    const uint64_t b = pl.starts[c], e = b + pl.counts[c];
    unsigned long long* w = scratch + word_base[c];
    // runs of 32 sorted values, one atomic per word a run touches
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

__global__ __launch_bounds__(256) void k_fc_bitmap_wide_write(const uint32_t* is_bitmap, const uint32_t* words,
                                                              const uint32_t* first,
                                                              const unsigned long long* word_base,
                                                              const unsigned long long* scratch,
                                                              const uint64_t* out_off, uint8_t* out,
                                                              uint64_t out_cap, uint32_t* err) {
    const uint32_t c = blockIdx.x;
    if (!is_bitmap[c] || words[c] <= FC_LDS_WORDS) return;
    if (out_off[c + 1] > out_cap) {  // not written; counted
        if (threadIdx.x == 0 && err) atomicAdd(err, 1u);
        return;
    }
    const uint32_t nw = words[c];
    uint8_t* o = out + out_off[c];
    if (threadIdx.x == 0) {
        const uint32_t h = nw | CODEC_MAP, f = first[c];
        for (int j = 0; j < 4; ++j) {
            o[j] = (uint8_t)(h >> (8 * j));
            o[4 + j] = (uint8_t)(f >> (8 * j));
        }
    }
    const uint8_t* src = reinterpret_cast<const uint8_t*>(scratch + word_base[c]);
    for (uint64_t i = threadIdx.x; i < (uint64_t)nw * 8; i += blockDim.x) o[8 + i] = src[i];
}

// words of the chunks k_fc_bitmap_wide_* handle, for the scan of their bases
__global__ void k_fc_wide_words(uint32_t n, const uint32_t* is_bitmap, const uint32_t* words, uint32_t* wide) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) wide[c] = (is_bitmap[c] && words[c] > FC_LDS_WORDS) ? words[c] : 0u;
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

int read_starts(std::vector<uint64_t>& starts, const uint64_t* d_starts, uint32_t nchunks, hipStream_t s) {
    if (int r = to_host(starts, d_starts, (size_t)nchunks + 1, s)) return r;
    for (uint32_t c = 0; c < nchunks; ++c)
        if (starts[c + 1] < starts[c]) return fail(PS_ERR_INVALID_ARGUMENT, "chunk starts must not decrease");
    return PS_OK;
}

// the decode kernel's LDS exceeds the default 64 KB limit
int svd_setup() {
    static bool lds_set = false;
    if (!lds_set) {
        PS_HIP_CHECK(hipFuncSetAttribute((const void*)k_svi_decode_chunk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)SVD_LDS));
        lds_set = true;
    }
    return PS_OK;
}

// Decodes the SVI chunks `sel` (each with a nonzero count) of a concatenated
// stream into values, one block per chunk.
int svi_decode_chunks(const uint8_t* d_in, const uint64_t* d_in_off, const std::vector<uint32_t>& sel,
                      const std::vector<uint32_t>& counts_all, const uint32_t* d_counts_all,
                      const uint64_t* d_value_starts, uint32_t* d_values, uint8_t* d_bounds, hipStream_t s) {
    (void)d_counts_all;
    const uint32_t n = (uint32_t)sel.size();
    if (n == 0) return PS_OK;
    for (uint32_t k = 0; k < n; ++k)
        if (counts_all[sel[k]] > MAX_CHUNK_VALUES)
            return fail(PS_ERR_INVALID_ARGUMENT, "stream chunk claims %u values", counts_all[sel[k]]);
    if (int r = svd_setup()) return r;
    DevBuf d_sel, d_err;
    if (int r = to_device(d_sel, sel, s)) return r;
    const std::vector<uint32_t> zero(1, 0u);
    if (int r = to_device(d_err, zero, s)) return r;
    hipLaunchKernelGGL(k_svi_decode_chunk, dim3(n), dim3(SVD_TPB), SVD_LDS, s, d_in, d_in_off, d_sel.as<uint32_t>(),
                       d_value_starts, d_bounds ? 1 : 0, d_values, d_bounds, d_err.as<uint32_t>(), ~0ull);
    PS_HIP_CHECK(hipGetLastError());
    std::vector<uint32_t> bad;
    if (int r = to_host(bad, d_err.p, 1, s)) return r;
    if (bad[0])
        return fail(PS_ERR_INVALID_ARGUMENT, "%u StreamVInt chunks do not end where their tuples end", bad[0]);
    return PS_OK;
}

}  // namespace

namespace {

// Plan tables of every chunk of `starts` in one device buffer (chunk starts,
// tuple bases, counts), copied from the host in one transfer.
int svi_plan_all(const std::vector<uint64_t>& starts, hipStream_t s, DevBuf& tab, SviPlan& pl) {
    const uint32_t n = (uint32_t)(starts.size() - 1);
    std::vector<uint64_t> h((size_t)2 * (n + 1) + (n + 1) / 2 + 1, 0);
    uint64_t* st = h.data();
    uint64_t* tb = st + n + 1;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(tb + n + 1);
    uint64_t t = 0, max_t = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t c0 = starts[k], c1 = starts[k + 1];
        if (c1 < c0) return fail(PS_ERR_INVALID_ARGUMENT, "chunk starts must not decrease");
        if (c1 - c0 > MAX_CHUNK_VALUES)
            return fail(PS_ERR_INVALID_ARGUMENT, "chunk of %llu values exceeds %u", (unsigned long long)(c1 - c0),
                        MAX_CHUNK_VALUES);
        st[k] = c0;
        cnt[k] = (uint32_t)(c1 - c0);
        tb[k] = t;
        const uint64_t nt = (uint64_t)((cnt[k] + 7) / 8 * 8) / 4;
        t += nt;
        max_t = std::max(max_t, nt);
    }
    st[n] = starts[n];
    tb[n] = t;
    if (int r = to_device(tab, h, s)) return r;
    const bool g2 = n > 0 && n <= 65535 && ((max_t + 255) / 256) * 256 * n <= 2 * t + 256ull * n;
    uint64_t* base = tab.as<uint64_t>();
    pl = SviPlan{base, base + n + 1, reinterpret_cast<const uint32_t*>(base + 2 * (n + 1)), n, t, (uint32_t)max_t,
                 g2 ? 1 : 0};
    return PS_OK;
}

// tuple sizes and the per-chunk payload sums of a plan
int svi_payload(const uint32_t* d_values, const SviPlan& pl, hipStream_t s, DevBuf& d_toff, DevBuf& d_pay,
                const uint32_t* skip = nullptr) {
    if (int r = d_toff.alloc((pl.ntuples + 1) * 4, s)) return r;
    if (int r = d_pay.alloc(((uint64_t)pl.nchunks + 1) * 8, s)) return r;
    if (pl.ntuples)
        hipLaunchKernelGGL(k_svi_tuple_sizes, svi_grid(pl), dim3(256), 0, s, d_values, pl, d_toff.as<uint32_t>(),
                           skip);
    if (pl.nchunks)
        hipLaunchKernelGGL(k_chunk_scan_u32<256>, dim3(pl.nchunks), dim3(256), 0, s, d_toff.as<uint32_t>(),
                           pl.tuple_base, d_pay.as<uint64_t>());
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

}  // namespace

int svi_encode_chunks_dev(const uint32_t* d_values, const uint8_t* d_bounds, const std::vector<uint64_t>& starts,
                          uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s) {
    DevBuf tab, d_toff, d_pay, d_bytes;
    SviPlan pl{};
    if (int r = svi_plan_all(starts, s, tab, pl)) return r;
    if (int r = svi_payload(d_values, pl, s, d_toff, d_pay)) return r;
    const uint32_t n = pl.nchunks;
    if (int r = d_bytes.alloc(((uint64_t)n + 1) * 8, s)) return r;
    if (n)
        hipLaunchKernelGGL(k_svi_chunk_bytes, dim3(blocks_for(n, 256)), dim3(256), 0, s, pl,
                           (const uint64_t*)d_pay.as<uint64_t>(), d_bounds ? 1 : 0, d_bytes.as<uint64_t>());
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl_u64(d_bytes.as<unsigned long long>(), n, (unsigned long long*)d_offsets, s)) return r;
    if (pl.ntuples)
        hipLaunchKernelGGL(k_svi_write, svi_grid(pl), dim3(256), 0, s, d_values, d_bounds, pl,
                           (const uint32_t*)d_toff.as<uint32_t>(), (const uint64_t*)d_offsets,
                           (const uint64_t*)d_pay.as<uint64_t>(), d_out, (const uint32_t*)nullptr, out_cap, d_err);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

int fc_encode_chunks_dev(const uint32_t* d_values, const std::vector<uint64_t>& starts, int with_check,
                         uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, uint32_t* d_err, hipStream_t s) {
    uint32_t* d_cap_err = d_err ? d_err + 1 : nullptr;
    const uint32_t n = (uint32_t)(starts.size() - 1);
    DevBuf tab, d_toff, d_pay, d_fl, d_bytes, d_wbase, d_scratch;
    SviPlan pl{};
    if (int r = svi_plan_all(starts, s, tab, pl)) return r;
    // first, last, ok, is_bitmap, words, wide words: six arrays of n
    if (int r = d_fl.alloc((size_t)6 * n * 4 + 4, s)) return r;
    uint32_t* first = d_fl.as<uint32_t>();
    uint32_t* last = first + n;
    uint32_t* ok = last + n;
    uint32_t* is_bm = ok + n;
    uint32_t* words = is_bm + n;
    uint32_t* wide = words + n;
    if (n) {
        hipLaunchKernelGGL(k_fc_inspect<1024>, dim3(n), dim3(1024), 0, s, d_values, pl.starts, first, last, ok);
        hipLaunchKernelGGL(k_fc_kind, dim3(blocks_for(n, 256)), dim3(256), 0, s, pl, (const uint32_t*)first,
                           (const uint32_t*)last, (const uint32_t*)ok, with_check, is_bm, words, d_err);
        PS_HIP_CHECK(hipGetLastError());
    }
    // StreamVInt sizes of the other chunks only
    if (int r = svi_payload(d_values, pl, s, d_toff, d_pay, is_bm)) return r;
    if (int r = d_bytes.alloc(((uint64_t)n + 1) * 8, s)) return r;
    if (n)
        hipLaunchKernelGGL(k_fc_bytes, dim3(blocks_for(n, 256)), dim3(256), 0, s, pl,
                           (const uint64_t*)d_pay.as<uint64_t>(), (const uint32_t*)is_bm, (const uint32_t*)words,
                           d_bytes.as<uint64_t>());
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl_u64(d_bytes.as<unsigned long long>(), n, (unsigned long long*)d_offsets, s)) return r;
    if (pl.ntuples)
        hipLaunchKernelGGL(k_svi_write, svi_grid(pl), dim3(256), 0, s, d_values, (const uint8_t*)nullptr, pl,
                           (const uint32_t*)d_toff.as<uint32_t>(), (const uint64_t*)d_offsets,
                           (const uint64_t*)d_pay.as<uint64_t>(), d_out, (const uint32_t*)is_bm, out_cap,
                           d_cap_err);
    if (n)
        hipLaunchKernelGGL(k_fc_bitmap_chunk, dim3(n), dim3(256), 0, s, d_values, pl, (const uint32_t*)is_bm,
                           (const uint32_t*)words, (const uint32_t*)first, (const uint64_t*)d_offsets, d_out, out_cap,
                           d_cap_err);
    PS_HIP_CHECK(hipGetLastError());
    // bitmap chunks wider than LDS: a bitmap has fewer than 10 bits per value,
    // so the chunks of more than 64 * FC_LDS_WORDS / 10 values bound the scratch
    uint64_t wide_cap = 0;
    for (uint32_t c = 0; c < n; ++c) {
        const uint64_t cnt = starts[c + 1] - starts[c];
        if (cnt * 10 > 64ull * FC_LDS_WORDS) wide_cap += cnt * 10 / 64 + 1;
    }
    if (wide_cap) {
        if (int r = d_wbase.alloc(((uint64_t)n + 1) * 8, s)) return r;
        if (int r = d_scratch.alloc(wide_cap * 8, s)) return r;
        if (int r = launch_fill(d_scratch.p, 0, wide_cap * 8, s)) return r;
        hipLaunchKernelGGL(k_fc_wide_words, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, (const uint32_t*)is_bm,
                           (const uint32_t*)words, wide);
        PS_HIP_CHECK(hipGetLastError());
        if (int r = scan_excl_u32(wide, n, d_wbase.as<unsigned long long>(), s)) return r;
        hipLaunchKernelGGL(k_fc_bitmap_wide_set, dim3(n), dim3(256), 0, s, d_values, pl, (const uint32_t*)is_bm,
                           (const uint32_t*)words, (const uint32_t*)first,
                           (const unsigned long long*)d_wbase.as<unsigned long long>(),
                           d_scratch.as<unsigned long long>());
        hipLaunchKernelGGL(k_fc_bitmap_wide_write, dim3(n), dim3(256), 0, s, (const uint32_t*)is_bm,
                           (const uint32_t*)words, (const uint32_t*)first,
                           (const unsigned long long*)d_wbase.as<unsigned long long>(),
                           (const unsigned long long*)d_scratch.as<unsigned long long>(), (const uint64_t*)d_offsets,
                           d_out, out_cap, d_cap_err);
        PS_HIP_CHECK(hipGetLastError());
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
    std::vector<uint64_t> starts;
    if (int r = read_starts(starts, d_starts, nchunks, s)) return r;
    // sizes, offsets and the write on the device; a chunk that would end past
    // out_cap is not written, and the call reports the capacity error
    DevBuf d_err;
    if (int r = d_err.alloc(8, s)) return r;
    if (int r = launch_fill(d_err.p, 0, 8, s)) return r;
    if (int r = svi_encode_chunks_dev(d_values, d_bounds, starts, d_out, out_cap, d_offsets, d_err.as<uint32_t>(), s))
        return r;
    uint64_t tail[2] = {0, 0};
    PS_HIP_CHECK(hipMemcpyAsync(&tail[0], d_offsets + nchunks, 8, hipMemcpyDeviceToHost, s));
    PS_HIP_CHECK(hipMemcpyAsync(&tail[1], d_err.p, 8, hipMemcpyDeviceToHost, s));
    if (int r = stream_wait(s)) return r;
    if (tail[1] || tail[0] > out_cap)
        return fail(PS_ERR_CAPACITY, "encoded size %llu exceeds output capacity %llu", (unsigned long long)tail[0],
                    (unsigned long long)out_cap);
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
    // counts, their prefix sums as the value starts, and every chunk decoded
    // with its limits checked on the device; one wait for the error counts
    DevBuf d_counts, d_err;
    if (int r = d_counts.alloc((size_t)nchunks * 4 + 4, s)) return r;
    if (int r = d_err.alloc(16, s)) return r;
    if (int r = launch_fill(d_err.p, 0, 16, s)) return r;
    if (nchunks)
        hipLaunchKernelGGL(k_read_counts, dim3(blocks_for(nchunks, 256)), dim3(256), 0, s, d_in, d_in_offsets, nchunks,
                           d_counts.as<uint32_t>(), (uint32_t*)nullptr);
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl_u32(d_counts.as<uint32_t>(), nchunks, (unsigned long long*)d_value_starts, s)) return r;
    if (nchunks) {
        if (int r = svd_setup()) return r;
        hipLaunchKernelGGL(k_svi_decode_chunk, dim3(nchunks), dim3(SVD_TPB), SVD_LDS, s, d_in, d_in_offsets,
                           (const uint32_t*)nullptr, (const uint64_t*)d_value_starts, with_bounds ? 1 : 0, d_values,
                           with_bounds ? d_bounds : nullptr, d_err.as<uint32_t>(), values_cap);
        PS_HIP_CHECK(hipGetLastError());
    }
    std::vector<uint32_t> bad;
    if (int r = to_host(bad, d_err.p, 3, s)) return r;
    if (bad[2]) return fail(PS_ERR_INVALID_ARGUMENT, "%u stream chunks claim more than %u values", bad[2], MAX_CHUNK_VALUES);
    if (bad[1])
        return fail(PS_ERR_CAPACITY, "decoded values exceed capacity %llu", (unsigned long long)values_cap);
    if (bad[0])
        return fail(PS_ERR_INVALID_ARGUMENT, "%u StreamVInt chunks do not end where their tuples end", bad[0]);
    return PS_OK;
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
    std::vector<uint64_t> starts;
    if (int r = read_starts(starts, d_starts, nchunks, s)) return r;
    // the choice, sizes, offsets and writes on the device; a chunk that would
    // end past out_cap is not written, and the call reports the capacity error
    DevBuf d_err;
    if (int r = d_err.alloc(8, s)) return r;
    if (int r = launch_fill(d_err.p, 0, 8, s)) return r;
    if (int r = fc_encode_chunks_dev(d_values, starts, with_check, d_out, out_cap, d_offsets, d_err.as<uint32_t>(), s))
        return r;
    uint64_t total = 0;
    uint32_t err[2] = {0, 0};
    PS_HIP_CHECK(hipMemcpyAsync(&total, d_offsets + nchunks, 8, hipMemcpyDeviceToHost, s));
    PS_HIP_CHECK(hipMemcpyAsync(err, d_err.p, 8, hipMemcpyDeviceToHost, s));
    if (int r = stream_wait(s)) return r;
    if (err[0])
        return fail(PS_ERR_INVALID_ARGUMENT, "%u chunks would be bitmaps but are not sorted; use the checked encoder",
                    err[0]);
    if (err[1] || total > out_cap)
        return fail(PS_ERR_CAPACITY, "encoded size %llu exceeds output capacity %llu", (unsigned long long)total,
                    (unsigned long long)out_cap);
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
    if (int r = stream_wait(s)) return r;
    return PS_OK;
    PS_API_END
}

}  // extern "C"
