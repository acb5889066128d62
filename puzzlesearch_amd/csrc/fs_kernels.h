// Kernels of one disk-segment step of the frontier search with 4-bit op flags,
// DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132).
//
// HBM layout of one segment (2^segment_bits states):
//   flags  u32[2^segment_bits / 8]: the IndexedArray<4> of the reference
//          (Common/BitArray.h:283-314): state i owns bits 4*(i%8)..4*(i%8)+3 of
//          word i/8, bit `op` set when a child arrived by the move labelled op
//   dirty  u8 per region of 2^16 states: set by any mark in the region, so the
//          scan passes skip untouched regions (the role of IndexedArray's
//          occupancy index)
//   odd    u64 bit array of the current frontier, only for puzzles with odd
//          cycles (the reference's FrontierArray)
#pragma once

#include "bfs_kernels.h"

namespace psk {

constexpr int FS_REGION_BITS = 16;                               // states per dirty region
constexpr uint32_t FS_REGION_WORDS = 1u << (FS_REGION_BITS - 3);  // u32 flag words per region
constexpr uint64_t FS_NONE = ~0ull;                              // empty cross-segment slot

struct FsMarkArgs {
    psd::PuzzleParams prm;
    const uint32_t* idx;  // frontier, in-segment ids
    const uint8_t* ops;   // their op bits
    uint64_t n;
    uint64_t base;  // segment << seg_bits
    uint32_t segment;
    int seg_bits;
    uint32_t* flags;
    uint8_t* dirty;
    unsigned long long* odd;       // NULL without odd cycles
    unsigned long long* counters;  // [0] in-segment children, [6] ids outside the segment
};

struct FsXsegArgs {
    const uint32_t* idx;  // cross-segment marks, per op
    const uint64_t* op_starts;  // ops + 1 entries (device)
    int ops;
    uint64_t seg_mask;
    uint32_t* flags;
    uint8_t* dirty;
};

struct FsScanArgs {
    const uint32_t* flags_c;
    uint32_t* flags;
    uint8_t* dirty;
    const unsigned long long* odd;
    uint64_t nwords;  // flag words of the segment
    int ops;
    unsigned long long* cnt_out;    // per region: entries written to the new frontier
    unsigned long long* counters;   // [1] all nonzero entries (the layer count), [2] written, [5] words scanned
    const unsigned long long* out_off;  // per region exclusive offsets (write pass)
    uint32_t* out_idx;
    uint8_t* out_ops;
};

struct FsCrossArgs {
    psd::PuzzleParams prm;
    const uint32_t* idx;
    const uint8_t* ops;
    uint64_t n;
    uint64_t base;
    uint32_t segment;
    uint32_t nseg;
    int seg_bits;
    unsigned long long* slots;     // n * SLOTS: (label * nseg + seg) << 32 | in-segment id, or FS_NONE
    unsigned long long* counters;  // [3] cross children, [4] children outside [0, nseg), [7] filled slots
    unsigned long long* block_count;  // per block of the cross pass: its cross children
};

// low bit of every nonzero nibble
__device__ __forceinline__ uint32_t nib_nonzero(uint32_t w) { return (w | (w >> 1) | (w >> 2) | (w >> 3)) & 0x11111111u; }

// low bit of every nibble whose low `ops` bits are all set (every move used)
__device__ __forceinline__ uint32_t nib_full(uint32_t w, int ops) {
    uint32_t f = w;
    if (ops > 1) f &= w >> 1;
    if (ops > 2) f &= w >> 2;
    if (ops > 3) f &= w >> 3;
    return f & 0x11111111u;
}

// bit k of b -> bit 4k
__device__ __forceinline__ uint32_t spread8(uint32_t b) {
    uint32_t x = b & 0xFFu;
    x = (x | (x << 12)) & 0x000F000Fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x;
}

// entries of flag word j that are counted: nonzero and not in the current frontier
__device__ __forceinline__ uint32_t fs_live(uint32_t w, const unsigned long long* odd, uint64_t j) {
    uint32_t nz = nib_nonzero(w);
    if (odd && nz) nz &= ~spread8(reinterpret_cast<const uint8_t*>(odd)[j]);
    return nz;
}

template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_fs_mark_in(FsMarkArgs a) {
    const uint64_t in_mask = (1ull << a.seg_bits) - 1ull;
    unsigned long long marks = 0, bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * TPB) {
        const uint32_t id = a.idx[i];
        if ((uint64_t)id > in_mask) {  // not an id of this segment: reported after the step
            ++bad;
            continue;
        }
        if (a.odd) atomicOr(&a.odd[id >> 6], 1ull << (id & 63));
        PZ::expand(a.base | id, (uint32_t)a.ops[i], a.prm, [&](int, uint64_t child, uint32_t label) {
            if ((child >> a.seg_bits) != a.segment) return;
            const uint64_t c = child & in_mask;
            gor(&a.flags[c >> 3], (1u << label) << ((c & 7) * 4));
            uint8_t* d = &a.dirty[c >> FS_REGION_BITS];
            if (!*d) *d = 1;
            ++marks;
        });
    }
    block_add3<TPB>(marks, bad, 0ull, &a.counters[0], &a.counters[6], nullptr);
}


// pass 1 of ScanBitsAndClear: per region, entries counted and entries written
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_count(FsScanArgs a) {
    const uint64_t r = blockIdx.x;
    const uint64_t w0 = r * FS_REGION_WORDS;
    const uint64_t rw = (a.nwords - w0 < FS_REGION_WORDS) ? a.nwords - w0 : FS_REGION_WORDS;
    unsigned long long all = 0, out = 0;
    if (a.dirty[r]) {
        // four words per load; a region is a multiple of 4 words unless it is the
        // whole (smaller) segment, whose tail is read word by word
        const uint64_t rw4 = rw & ~3ull;
        const uint4* f4 = reinterpret_cast<const uint4*>(a.flags_c + w0);
        for (uint64_t j4 = threadIdx.x; j4 * 4 < rw4; j4 += TPB) {
            const uint4 v = f4[j4];
            const uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if (!ws[e]) continue;
                const uint32_t live = fs_live(ws[e], a.odd, w0 + j4 * 4 + e);
                all += __popc(live);
                out += __popc(live & ~nib_full(ws[e], a.ops));
            }
        }
        for (uint64_t j = rw4 + threadIdx.x; j < rw; j += TPB) {
            const uint32_t w = a.flags_c[w0 + j];
            if (!w) continue;
            const uint32_t live = fs_live(w, a.odd, w0 + j);
            all += __popc(live);
            out += __popc(live & ~nib_full(w, a.ops));
        }
    }
    all = wave_sum(all);
    out = wave_sum(out);
    __shared__ unsigned long long red[2][TPB / 64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[0][wave] = all;
        red[1][wave] = out;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sa = 0, so = 0;
        for (int w = 0; w < TPB / 64; ++w) {
            sa += red[0][w];
            so += red[1][w];
        }
        a.cnt_out[r] = so;
        if (a.dirty[r]) atomicAdd(&a.counters[5], (unsigned long long)rw);
        if (sa) atomicAdd(&a.counters[1], sa);
        if (so) atomicAdd(&a.counters[2], so);
    }
}

// pass 2: ordered write of (id, op bits) for entries with an unused move, and
// the clear of the region (flags words and dirty byte)
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_write(FsScanArgs a) {
    const uint64_t r = blockIdx.x;
    if (!a.dirty[r]) return;
    const uint64_t w0 = r * FS_REGION_WORDS;
    const uint64_t rw = (a.nwords - w0 < FS_REGION_WORDS) ? a.nwords - w0 : FS_REGION_WORDS;
    __shared__ uint32_t wave_tot[TPB / 64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    uint64_t running = a.out_off[r];
    // thread t of iteration `it` owns words it*4*TPB + 4t .. +3: (iteration, thread,
    // word) order is index order, so one block scan per iteration keeps the output sorted
    for (uint64_t it = 0; it * 4 * TPB < rw; ++it) {
        const uint64_t j0 = it * 4 * TPB + (uint64_t)threadIdx.x * 4;
        uint32_t ws[4] = {0u, 0u, 0u, 0u}, keep[4] = {0u, 0u, 0u, 0u};
        uint32_t c = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint64_t j = j0 + e;
            ws[e] = (j < rw) ? a.flags[w0 + j] : 0u;
            if (ws[e]) keep[e] = fs_live(ws[e], a.odd, w0 + j) & ~nib_full(ws[e], a.ops);
            c += __popc(keep[e]);
        }
        const uint32_t incl = wave_incl_scan(c, lane);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int k = 0; k < TPB / 64; ++k) {
            const uint32_t t = wave_tot[k];
            before += (k < wave) ? t : 0u;
            total += t;
        }
        uint64_t pos = running + before + incl - c;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint32_t kb = keep[e];
            while (kb) {
                const int b = __ffs(kb) - 1;  // bit 4k of nibble k
                kb &= kb - 1;
                a.out_idx[pos] = (uint32_t)((w0 + j0 + e) * 8 + (uint64_t)(b >> 2));
                a.out_ops[pos] = (uint8_t)((ws[e] >> b) & 15u);
                ++pos;
            }
            if (ws[e]) a.flags[w0 + j0 + e] = 0u;
        }
        running += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) a.dirty[r] = 0;
}


// children of the written entries that leave the segment, one slot per move
struct CrossSink {
    unsigned long long* out;
    uint64_t segment, nseg, in_mask;
    int seg_bits;
    unsigned long long nx, bad;
    __device__ __forceinline__ void operator()(int slot, uint64_t child, uint32_t label) {
        const uint64_t s = child >> seg_bits;
        if (s == segment) return;
        if (s >= nseg) {
            ++bad;
            return;
        }
        out[slot] = (((uint64_t)label * nseg + s) << 32) | (child & in_mask);
        ++nx;
    }
};

constexpr uint32_t FS_CROSS_PER_BLOCK = 4096;  // parents per block of the cross pass

// Block b expands parents [b*4096, (b+1)*4096) and records how many of their
// slots it filled, so the compaction can place each block's children in
// (parent, slot) order.  The count is of the slots themselves, read back after
// the expansion, so the compaction writes exactly sum(block_count) entries;
// counters[7] receives the same total for the host to compare with the
// children counted in counters[3].
template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_fs_cross(FsCrossArgs a) {
    CrossSink sink{nullptr, a.segment, a.nseg, (1ull << a.seg_bits) - 1ull, a.seg_bits, 0ull, 0ull};
    const uint64_t p0 = (uint64_t)blockIdx.x * FS_CROSS_PER_BLOCK;
    const uint64_t p1 = (p0 + FS_CROSS_PER_BLOCK < a.n) ? p0 + FS_CROSS_PER_BLOCK : a.n;
    unsigned long long filled = 0;
    for (uint64_t i = p0 + threadIdx.x; i < p1; i += TPB) {
        sink.out = a.slots + i * PZ::SLOTS;
#pragma unroll
        for (int k = 0; k < PZ::SLOTS; ++k) sink.out[k] = FS_NONE;
        PZ::expand(a.base | a.idx[i], (uint32_t)a.ops[i], a.prm, sink);
#pragma unroll
        for (int k = 0; k < PZ::SLOTS; ++k) filled += sink.out[k] != FS_NONE ? 1ull : 0ull;
    }
    block_add3<TPB>(sink.nx, sink.bad, filled, &a.counters[3], &a.counters[4], &a.counters[7]);
    __shared__ unsigned long long red[TPB / 64];
    unsigned long long v = wave_sum(filled);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < TPB / 64; ++w) t += red[w];
        a.block_count[blockIdx.x] = t;
    }
}

// Ordered compaction of the cross slots: block b copies the non-empty slots of
// its parents to out[block_off[b] ...] in slot order.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_compact(const unsigned long long* __restrict__ slots, uint64_t nslots,
                                                    uint32_t slots_per_block,
                                                    const unsigned long long* __restrict__ block_off,
                                                    unsigned long long* __restrict__ out) {
    __shared__ uint32_t wave_tot[TPB / 64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint64_t s0 = (uint64_t)blockIdx.x * slots_per_block;
    const uint64_t s1 = (s0 + slots_per_block < nslots) ? s0 + slots_per_block : nslots;
    uint64_t running = block_off[blockIdx.x];
    for (uint64_t base = s0; base < s1; base += TPB) {
        const uint64_t i = base + threadIdx.x;
        const unsigned long long x = (i < s1) ? slots[i] : FS_NONE;
        const uint32_t c = x != FS_NONE ? 1u : 0u;
        const uint32_t incl = wave_incl_scan(c, lane);
        if (lane == 63) wave_tot[wave] = incl;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int k = 0; k < TPB / 64; ++k) {
            const uint32_t t = wave_tot[k];
            before += (k < wave) ? t : 0u;
            total += t;
        }
        if (c) out[running + before + incl - 1] = x;
        running += total;
        __syncthreads();
    }
}


}  // namespace psk
